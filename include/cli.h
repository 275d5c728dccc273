// SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0
// The command-line contract of both tools (reference src/cli.h): flag parsing, the weather / time rules after parsing,
// and the output manifests. Return codes are the exit codes: 0 ok, 1 usage error, 2 runtime error.
#pragma once
#include <functional>
#include <string>
#include <vector>

#include "daylight.h"
#include "views.h"

namespace daylight::cli {

extern const char* const kUsage;

struct CommonArgs {
  std::vector<std::string> sensors, context;
  std::string out;
  bool bench = false, help = false;
  size_t benchSens = 1000000, benchTris = 10000000;
};

struct Args : CommonArgs {
  Options opt;
  std::string epw, timeStr, fromStr, toStr;
  bool clearSky = false;
  double turbidity = 3, altitude = 0;
};

int parse_common(int argc, char** argv, CommonArgs& a, RunOptions& o, std::string& err,
                 const std::function<int(int& i, const std::string& flag)>& tool);
int parse(int argc, char** argv, Args& a, std::string& err);
int prepare(Args& a, Weather& w, bool& haveWeather, std::string& err);
std::string manifest_json(const Result& r, const std::string& extraFields = "");

}  // namespace daylight::cli

namespace daylight::views::cli {

extern const char* const kUsage;

struct Args : daylight::cli::CommonArgs {
  Options opt;
};

int parse(int argc, char** argv, Args& a, std::string& err);
std::string manifest_json(const Result& r, const Options& opt, const Field& f);

}  // namespace daylight::views::cli
