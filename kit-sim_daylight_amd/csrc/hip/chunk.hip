// SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0
// The device work of one daylight chunk as a HIP graph: the visibility trace, the reduce of every quantity, and with
// a series its contraction (the ground view factor taken from the reduce's last row on the device), launched on the
// library's stream as one graph. The reference's chunk loop launches and waits kernel by kernel and sizes TDR-safe
// sub-launches from the last one's time (reference src/run.cpp:625-688, src/gpu.cu:505-531); here a chunk is one
// launch of each kernel over all its sensors, and the host meets the device once per chunk.
//
// The graph is captured on the first chunk of a shape and replayed while every kernel argument stays the same: the
// buffers (which only grow), the sizes, the flags and the chosen kernels make its key. Uploads (weights, series inputs)
// run before the launch on the same stream; nothing in the graph sets memory from the host (no memset nodes).
#include <cstdlib>
#include <cstring>
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
                        g->out.p, g->seriesIn.p, g->seriesE.p, g->seriesWork.p, g->skyS.p, g->skyG.p, g->skySdc.p, g->dc.p, g->stackOvf.p, g->tileNext.p};
  std::memcpy(k.p, ptrs, sizeof ptrs);
  const dlg_series_plan& q = g->seriesPlan;
  const int64_t vals[] = {s.nTris, pad_bits(s.pad), s.nNodes, g->nDir, g->nWords, g->nc, g->i0, int64_t(rayFlags), stack, g->bytes,
                          g->nSlots, nQ, int64_t(mask), series ? 1 : 0, q.method, q.groundRow, q.reflect, q.T, int64_t(q.nc) << 32 | q.nSky,
                          int64_t(q.nBins) << 32 | q.sun0, int64_t(q.oRow ^ (q.oW << 16) ^ (q.oG << 32)), int64_t(q.oCol ^ (q.oSw << 32)), g->dcBins,
                          q.shape, g->traceGrid};
  std::memcpy(k.n, vals, sizeof vals);
  return k;
}

}  // namespace

struct dlg_graph_cache {
  GraphKey key;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t ev[4] = {};   // recorded by the graph around its stages (external event nodes): trace, reduce, series
  hipEvent_t launch[2] = {};   // recorded around each launch, outside the graph
  bool staged = true;          // whether the stage events give times (else the launch's time is the trace's)
  void drop() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
  ~dlg_graph_cache() {
    drop();
    for (hipEvent_t e : ev) if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : launch) if (e) (void)hipEventDestroy(e);
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

// Captures the chunk's kernels on g's stream into c (false: the capture failed; the stream is out of capture again).
bool capture(dlg_gpu* g, dlg_graph_cache& c, const dlg_chunk_in* in, int stack, uint32_t mask, bool series) {
  if (hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return false;
  // an event recorded under capture is only an ordering point unless it is external: then it is a node that records
  const auto mark = [&](hipEvent_t e) { return hipEventRecordWithFlags(e, g->stream, hipEventRecordExternal) == hipSuccess; };
  bool ok = mark(c.ev[0]);
  dlg_trace_enqueue(g, in->rayFlags, stack, nullptr);
  ok = mark(c.ev[1]) && ok;
  dlg_reduce_enqueue(g, uint32_t(in->nc), in->nQ, mask);
  ok = mark(c.ev[2]) && ok;
  if (series) dlg_series_enqueue(g);
  ok = mark(c.ev[3]) && ok;
  ok = hipGetLastError() == hipSuccess && ok;
  hipGraph_t graph = nullptr;
  ok = hipStreamEndCapture(g->stream, &graph) == hipSuccess && ok && graph;
  if (ok) ok = hipGraphInstantiate(&c.exec, graph, nullptr, nullptr, 0) == hipSuccess;
  if (!ok) {
    if (graph) (void)hipGraphDestroy(graph);
    c.exec = nullptr;
    (void)hipGetLastError();
    return false;
  }
  c.graph = graph;
  return true;
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
    for (hipEvent_t& e : g->graph->launch) DLG_CHECK(hipEventCreate(&e));
  }
  dlg_graph_cache& c = *g->graph;
  const GraphKey key = key_of(g, in->rayFlags, stack, in->nQ, mask, series);
  if (!c.exec || std::memcmp(&key, &c.key, sizeof key) != 0) {
    c.drop();
    if (!capture(g, c, in, stack, mask, series)) {   // this context runs its chunks eagerly from now on
      g->graphBroken = 1;
      return chunk_eager(g, in, series, sums, E, err, errLen);
    }
    c.key = key;
    g->stats.graphCaptures += 1;
  }
  DLG_CHECK(hipEventRecord(c.launch[0], g->stream));
  DLG_CHECK(hipGraphLaunch(c.exec, g->stream));
  DLG_CHECK(hipEventRecord(c.launch[1], g->stream));
  g->stats.graphLaunches += 1;
  DLG_CHECK(hipMemcpyAsync(sums, g->out.p, size_t(in->nQ) * size_t(nc) * 8, hipMemcpyDeviceToHost, g->stream));
  if (series && E) DLG_CHECK(hipMemcpyAsync(E, g->seriesE.p, size_t(nc) * size_t(in->series->T) * 4, hipMemcpyDeviceToHost, g->stream));
  DLG_CHECK(hipStreamSynchronize(g->stream));
  float ms[3] = {0, 0, 0};
  for (int k = 0; k < 3 && c.staged; ++k) c.staged = hipEventElapsedTime(&ms[k], c.ev[k], c.ev[k + 1]) == hipSuccess;
  if (!c.staged) {   // a runtime without event nodes: the whole launch counts as the trace's
    (void)hipGetLastError();
    ms[1] = ms[2] = 0;
    DLG_CHECK(hipEventElapsedTime(&ms[0], c.launch[0], c.launch[1]));
  }
  if (dlg_trace_account(g, ms[0], err, errLen)) return 1;
  g->stats.reduceMs += ms[1];
  g->stats.lastReduceMs = ms[1];
  if (series) {
    g->stats.seriesMs += ms[2];
    g->stats.lastSeriesMs = ms[2];
  }
  return 0;
}
