// tune.cpp -- autotuner of the launch configuration (C-ABI mr_ctx_autotune).
//
// For a representative batch of frames it times every built combination of
// tile size, raster workgroup size and split size of crowded tiles,
// with HIP events on the context's stream, keeps the fastest, and sizes the
// bin list from the measured (tile, triangle) pairs per triangle.  Results
// can be kept in a small text cache keyed by the GPU and the batch's shape
// (frames, triangles, pixels, each rounded to a power of two), so a service
// tunes once per GPU and workload class.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "minirender_amd.h"
#include "context.h"

using namespace mr;

namespace {

int ilog2ceil(uint64_t v) {
  int r = 0;
  while ((1ull << r) < v && r < 63) r++;
  return r;
}

std::string tune_key(mr_ctx* c, const mr_frame* frames, int n) {
  hipDeviceProp_t prop;
  std::string arch = "unknown";
  if (hipGetDeviceProperties(&prop, c->device) == hipSuccess) arch = prop.gcnArchName;
  for (char& ch : arch)
    if (ch == ' ') ch = '_';
  uint64_t tris = 0, pix = 0;
  for (int i = 0; i < n; i++) {
    pix += (uint64_t)frames[i].width * frames[i].height;
    for (int d = 0; d < frames[i].ndraws; d++) {
      const int h = frames[i].draws[d].mesh;
      if (h >= 0 && h < (int)c->meshes.size()) tris += (uint64_t)c->meshes[h].ntri;
    }
  }
  char buf[256];
  snprintf(buf, sizeof(buf), "%s/cus%d/n%d/t%d/p%d", arch.c_str(), c->num_cus,
           ilog2ceil((uint64_t)n), ilog2ceil(tris + 1), ilog2ceil(pix + 1));
  return buf;
}

bool cache_lookup(const char* path, const std::string& key, mr_tuning& out) {
  if (!path) return false;
  FILE* f = fopen(path, "r");
  if (!f) return false;
  char line[512];
  bool found = false;
  while (fgets(line, sizeof(line), f)) {
    char k[256];
    mr_tuning t{};
    if (sscanf(line, "%255s %d %d %f %d %lf", k, &t.tile, &t.raster_threads, &t.bin_factor,
               &t.frag_size, &t.ms) == 6 &&
        key == k) {
      out = t;  // the last entry for a key wins
      found = true;
    }
  }
  fclose(f);
  return found;
}

void cache_store(const char* path, const std::string& key, const mr_tuning& t) {
  if (!path) return;
  FILE* f = fopen(path, "a");
  if (!f) return;
  fprintf(f, "%s %d %d %.3f %d %.6f\n", key.c_str(), t.tile, t.raster_threads, t.bin_factor,
          t.frag_size, t.ms);
  fclose(f);
}

}  // namespace

extern "C" {

int mr_ctx_autotune(mr_ctx* c, const mr_frame* frames, int32_t n, const mr_outputs* outs,
                    int32_t reps, const char* cache_path, mr_tuning* best) {
  if (!c || !frames || !outs || n < 1 || reps < 1 || reps > 1000) return MR_ERR_ARG;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  if (hipSetDevice(c->device) != hipSuccess) return MR_ERR_HIP;
  const std::string key = tune_key(c, frames, n);
  mr_tuning pick{};
  if (cache_lookup(cache_path, key, pick) && mr_ctx_set_tile(c, pick.tile, pick.tile) == MR_OK &&
      mr_ctx_set_raster_threads(c, pick.raster_threads) == MR_OK &&
      mr_ctx_set_bin_factor(c, pick.bin_factor) == MR_OK && pick.frag_size >= 64) {
    c->frag_size = (uint32_t)pick.frag_size;
    pick.from_cache = 1;
    if (best) *best = pick;
    return MR_OK;
  }
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return MR_ERR_HIP;
  if (hipEventCreate(&e1) != hipSuccess) {
    hipEventDestroy(e0);
    return MR_ERR_HIP;
  }
  const bool prof = c->profiling;
  c->profiling = false;
  // Generous bin list while measuring, so no candidate pays for a re-run.
  const float bf_saved = c->bin_factor;
  c->bin_factor = std::max(bf_saved, 4.0f);
  const uint32_t frag_saved = c->frag_size;
  struct Cand { int tile, nt, frag; };
  std::vector<Cand> cands;
  const int tiles[3] = {8, 16, 32};
  const int nts[4] = {64, 128, 256, 512};
  for (int t : tiles)
    for (int nt : nts)
      if (raster_threads_supported(t, nt))
        for (int frag : {256, 1024, 4096}) cands.push_back({t, nt, frag});
  int rc = MR_OK;
  double best_ms = 1e300;
  Cand best_c = {16, 256, 1024};
  int64_t pairs = 0, tris = 0;
  for (const Cand& cd : cands) {
    if ((rc = mr_ctx_set_tile(c, cd.tile, cd.tile))) break;
    if ((rc = mr_ctx_set_raster_threads(c, cd.nt))) break;
    c->frag_size = (uint32_t)cd.frag;
    if ((rc = render_frames(c, frames, n, outs))) break;  // warm: scratch sized
    std::vector<double> ts;
    for (int r = 0; r < reps && rc == MR_OK; r++) {
      hipEventRecord(e0, c->main.stream);
      rc = render_frames(c, frames, n, outs);
      hipEventRecord(e1, c->main.stream);
      hipEventSynchronize(e1);
      float ms = 0;
      hipEventElapsedTime(&ms, e0, e1);
      ts.push_back(ms);
    }
    if (rc) break;
    std::sort(ts.begin(), ts.end());
    const double med = ts[ts.size() / 2];
    if (med < best_ms) {
      best_ms = med;
      best_c = cd;
      pairs = c->stats.bin_entries;
      tris = c->stats.triangles_in;
    }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  c->profiling = prof;
  if (rc) {
    c->bin_factor = bf_saved;
    c->frag_size = frag_saved;
    return rc;
  }
  mr_ctx_set_tile(c, best_c.tile, best_c.tile);
  mr_ctx_set_raster_threads(c, best_c.nt);
  c->frag_size = (uint32_t)best_c.frag;
  // Bin capacity: the measured pairs per triangle with 25% headroom.
  float bf = tris > 0 ? (float)((double)pairs / (double)tris * 1.25) : bf_saved;
  bf = std::min(std::max(bf, 1.0f), 1024.0f);
  c->bin_factor = bf;
  pick.tile = best_c.tile;
  pick.raster_threads = best_c.nt;
  pick.frag_size = best_c.frag;
  pick.bin_factor = bf;
  pick.ms = best_ms;
  pick.from_cache = 0;
  pick.candidates = (int32_t)cands.size();
  cache_store(cache_path, key, pick);
  if (best) *best = pick;
  return MR_OK;
}

}  // extern "C"
