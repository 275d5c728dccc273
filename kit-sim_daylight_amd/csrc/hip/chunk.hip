// SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0
// The device work of one daylight chunk as HIP graphs: the visibility trace, the reduce of every quantity, and with a
// series its contraction (the ground view factor taken from the reduce's last row on the device), one graph per stage,
// launched back to back on the library's stream with an event between stages (their times), and the host waiting once.
// The reference's chunk loop launches and waits kernel by kernel and sizes TDR-safe sub-launches from the last one's
// time (reference src/run.cpp:625-688, src/gpu.cu:505-531); here a chunk is one launch of each kernel over all its
// sensors, and the host meets the device once per chunk.
//
// The graphs are captured on the first chunk of a shape and replayed while every kernel argument stays the same: the
// buffers (which only grow), the sizes, the flags and the chosen kernels make their key. Uploads (weights, series
// inputs) run before the launch on the same stream; nothing in a graph sets memory from the host (no memset nodes).
// The stage events are recorded on the stream between graph launches, not inside a graph (this runtime refuses event
// record nodes under capture).
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "gpu_internal.h"

namespace {

// every value a kernel of the chunk reads as an argument
struct GraphKey {
  const void* p[24];
  int64_t n[32];
};

int64_t pad_bits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return int64_t(u);
}

GraphKey key_of(const dlg_gpu* g, uint32_t rayFlags, int stack, int32_t nQ, uint32_t mask, bool series) {
  GraphKey k;
  std::memset(&k, 0, sizeof k);
  const dlg_scene& s = g->scene;
  const void* ptrs[] = {s.wide.p, s.leaves.p, g->P.p, g->N.p, g->D.p, g->dirTab.p, g->bits.p, g->slots.p, g->traceCount.p, g->scratch.p,
                        g->out.p, g->seriesIn.p, g->seriesE.p, g->seriesWork.p, g->skyS.p, g->skyG.p, g->skySdc.p, g->dc.p, g->stackOvf.p, g->tileNext.p,
                        g->tileOrder.p};
  std::memcpy(k.p, ptrs, sizeof ptrs);
  const dlg_series_plan& q = g->seriesPlan;
  const int64_t vals[] = {s.nTris, pad_bits(s.pad), s.nNodes, g->nDir, g->nWords, g->nc, g->i0, int64_t(rayFlags), stack, g->bytes,
                          g->nSlots, nQ, int64_t(mask), series ? 1 : 0, q.method, q.groundRow, q.reflect, q.T, int64_t(q.nc) << 32 | q.nSky,
                          int64_t(q.nBins) << 32 | q.sun0, int64_t(q.oRow ^ (q.oW << 16) ^ (q.oG << 32)), int64_t(q.oCol ^ (q.oSw << 32)), g->dcBins,
                          q.shape, g->traceGrid, int64_t(q.oMx)};
  std::memcpy(k.n, vals, sizeof vals);
  return k;
}

}  // namespace

struct dlg_graph_cache {
  GraphKey key;
  hipGraph_t graph[3] = {};        // the stages: trace, reduce, series
  hipGraphExec_t exec[3] = {};
  hipEvent_t ev[4] = {};           // recorded on the stream around the stage launches
  void drop() {
    for (int k = 0; k < 3; ++k) {
      if (exec[k]) (void)hipGraphExecDestroy(exec[k]);
      if (graph[k]) (void)hipGraphDestroy(graph[k]);
      exec[k] = nullptr;
      graph[k] = nullptr;
    }
  }
  ~dlg_graph_cache() {
    drop();
    for (hipEvent_t e : ev) if (e) (void)hipEventDestroy(e);
  }
};

void dlg_graph_free(dlg_gpu* g) {
  delete g->graph;
  g->graph = nullptr;
}

namespace {

// the chunk's stages one by one, each launched and waited for (the counting variant, DAYLIGHT_GRAPH=off, or a device
// where capture fails): the same kernels, so the same results
int32_t chunk_eager(dlg_gpu* g, const dlg_chunk_in* in, bool series, double* sums, float* E, char* err, int32_t errLen) {
  if (dlg_trace(g, in->i0, in->nc, in->rayFlags, err, errLen) || dlg_reduce(g, in->nc, in->nQ, in->W, in->useCos, sums, err, errLen)) return 1;
  if (!series) return 0;
  std::vector<float> G(static_cast<size_t>(in->nc));
  const double* ground = sums + size_t(in->nQ - 1) * size_t(in->nc);
  for (size_t i = 0; i < G.size(); ++i) G[i] = float(ground[i]);
  dlg_series_in s = *in->series;
  s.G = G.data();
  return dlg_series(g, in->nc, &s, in->method, E, err, errLen);
}

bool graphs_off() {
  const char* e = std::getenv("DAYLIGHT_GRAPH");
  return e && (!std::strcmp(e, "off") || !std::strcmp(e, "0"));
}

// Captures one stage's launches (enqueue) on g's stream into graph k of c (false: the capture failed, why in `why`;
// the stream is out of capture again).
template <class F>
bool capture_stage(dlg_gpu* g, dlg_graph_cache& c, int k, const char* name, F&& enqueue, std::string& why) {
  (void)hipGetLastError();   // an error left by an earlier call is not the capture's
  hipError_t r = hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal);
  if (r != hipSuccess) { why = std::string(name) + ": hipStreamBeginCapture: " + hipGetErrorString(r); return false; }
  enqueue();
  const hipError_t launched = hipGetLastError();
  hipGraph_t graph = nullptr;
  r = hipStreamEndCapture(g->stream, &graph);
  if (launched != hipSuccess || r != hipSuccess || !graph) {
    why = std::string(name) + ": " + hipGetErrorString(launched != hipSuccess ? launched : r);
    if (graph) (void)hipGraphDestroy(graph);
    (void)hipGetLastError();
    return false;
  }
  r = hipGraphInstantiate(&c.exec[k], graph, nullptr, nullptr, 0);
  if (r != hipSuccess) {
    why = std::string(name) + ": hipGraphInstantiate: " + hipGetErrorString(r);
    (void)hipGraphDestroy(graph);
    c.exec[k] = nullptr;
    (void)hipGetLastError();
    return false;
  }
  c.graph[k] = graph;
  return true;
}

bool capture(dlg_gpu* g, dlg_graph_cache& c, const dlg_chunk_in* in, int stack, uint32_t mask, bool series, std::string& why) {
  return capture_stage(g, c, 0, "trace", [&] { dlg_trace_enqueue(g, in->rayFlags, stack, nullptr); }, why) &&
         capture_stage(g, c, 1, "reduce", [&] { dlg_reduce_enqueue(g, uint32_t(in->nc), in->nQ, mask); }, why) &&
         (!series || capture_stage(g, c, 2, "series", [&] { dlg_series_enqueue(g); }, why));
}

}  // namespace

extern "C" int32_t dlg_chunk(dlg_gpu* g, const dlg_chunk_in* in, double* sums, float* E, char* err, int32_t errLen) {
  const int64_t nc = in->nc;
  const bool series = in->series && in->series->T > 0;
  if (in->nQ < 1 || in->nQ > 8) { dlg_fail(err, errLen, "dlg_chunk: 1 to 8 rows of weights"); return 1; }
  if (nc == 0) return 0;
  if (g->countFetches || graphs_off() || g->graphBroken) return chunk_eager(g, in, series, sums, E, err, errLen);
  int stack = 0;
  if (dlg_trace_prepare(g, in->i0, nc, in->rayFlags, stack, err, errLen)) return 1;
  if (g->nWords == 0) { dlg_fail(err, errLen, "dlg_chunk: no directions"); return 1; }
  uint32_t mask = 0;
  if (dlg_reduce_prepare(g, nc, in->nQ, in->W, in->useCos, mask, err, errLen)) return 1;
  if (series && dlg_series_prepare(g, nc, in->series, in->method, in->nQ - 1, err, errLen)) return 1;
  if (series && !dlg_series_tuned(g)) return chunk_eager(g, in, series, sums, E, err, errLen);   // dlg_series tunes it
  if (!g->graph) {
    g->graph = new dlg_graph_cache;
    for (hipEvent_t& e : g->graph->ev) DLG_CHECK(hipEventCreate(&e));
  }
  dlg_graph_cache& c = *g->graph;
  const GraphKey key = key_of(g, in->rayFlags, stack, in->nQ, mask, series);
  if (!c.exec[0] || std::memcmp(&key, &c.key, sizeof key) != 0) {
    c.drop();
    std::string why;
    if (!capture(g, c, in, stack, mask, series, why)) {   // this context runs its chunks eagerly from now on
      c.drop();
      g->graphBroken = 1;
      std::fprintf(stderr, "daylight: chunk graph capture failed (%s); this device context runs its chunks stage by stage\n", why.c_str());
      return chunk_eager(g, in, series, sums, E, err, errLen);
    }
    c.key = key;
    g->stats.graphCaptures += 1;
  }
  DLG_CHECK(hipEventRecord(c.ev[0], g->stream));
  DLG_CHECK(hipGraphLaunch(c.exec[0], g->stream));
  DLG_CHECK(hipEventRecord(c.ev[1], g->stream));
  DLG_CHECK(hipGraphLaunch(c.exec[1], g->stream));
  DLG_CHECK(hipEventRecord(c.ev[2], g->stream));
  if (series) DLG_CHECK(hipGraphLaunch(c.exec[2], g->stream));
  DLG_CHECK(hipEventRecord(c.ev[3], g->stream));
  g->stats.graphLaunches += 1;
  DLG_CHECK(hipMemcpyAsync(sums, g->out.p, size_t(in->nQ) * size_t(nc) * 8, hipMemcpyDeviceToHost, g->stream));
  if (series && E) DLG_CHECK(hipMemcpyAsync(E, g->seriesE.p, size_t(nc) * size_t(in->series->T) * 4, hipMemcpyDeviceToHost, g->stream));
  DLG_CHECK(hipStreamSynchronize(g->stream));
  float ms[3] = {0, 0, 0};
  for (int k = 0; k < 3; ++k) DLG_CHECK(hipEventElapsedTime(&ms[k], c.ev[k], c.ev[k + 1]));
  if (dlg_trace_account(g, ms[0], err, errLen)) return 1;
  g->stats.reduceMs += ms[1];
  g->stats.lastReduceMs = ms[1];
  if (series) {
    g->stats.seriesMs += ms[2];
    g->stats.lastSeriesMs = ms[2];
  }
  return 0;
}
