// Ragged batches: thousands of transforms of different 13-smooth sizes, packed back to
// back, in one persistent launch.
//
// The host groups the transforms by size ("classes"; one plan per class, each plan the
// same Stockham schedule the uniform path uses) and cuts the work into tiles of at most
// LDS_ELEMS elements:
//   * a single-stage class (N <= LDS_ELEMS): a tile is G = LDS_ELEMS / N transforms of
//     that size, listed by their element offsets;
//   * a two-stage class (N = P * Q): stage-A tiles of G_A columns, stage-B tiles of G_B
//     residue classes, as in the uniform path; stage A writes its intermediate into the
//     output buffer and stage B transforms it there in place.
// Tiles are queued largest class first (longest-processing-time order); workgroups take
// them from the queue with an agent-scope atomic counter until it runs out.
//
// Default ("split") schedule: one persistent launch runs every stage-A tile and every
// single-stage tile, a second runs every stage-B tile; the kernel boundary orders the
// stages. Alternative ("one", MFFT_RAGGED_MODE=one): a single launch in which the stage-B
// tiles of a transform are queued a few hundred tiles after its stage-A tiles and wait
// (bounded spin) until the transform's stage-A counter, raised with release semantics after
// each stage-A tile, reaches the number of stage-A tiles. Stage-A tiles are always queued
// before the stage-B tiles of their transform and never wait, so every wait ends; if a spin
// bound is ever hit the launch records an error flag and still exits.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <string>
#include <vector>

#include "../../include/mfft.h"
#include "mfft_pass.h"

namespace mfft {

struct RTile {
  int32_t cls;
  int32_t stage;  // 0: single stage or stage A, 1: stage B
  int32_t count;  // sub-problems (transforms, columns or residues) in the tile
  int32_t dep;    // two-stage: the transform's counter slot; -1 otherwise
  int64_t base;   // single stage: index into offs[]; two stage: element offset of the transform
  int64_t sub0;   // two stage: first column / residue
};

constexpr unsigned kSpinLimit = 1u << 26;

template <bool FWD>
__global__ __launch_bounds__(WG_THREADS) void ragged_kernel(
    const double2 *in, double2 *out, const double2 *__restrict__ tw,
    const RClass *__restrict__ classes, const RTile *__restrict__ tiles, int ntiles,
    const int64_t *__restrict__ offs, unsigned *head, unsigned *err,
    unsigned *deps /* stage-A tiles done, per two-stage transform */,
    const unsigned *__restrict__ need,
    int wait /* 0: the stages run as separate launches, no cross-workgroup waits */) {
  extern __shared__ double2 lds[];
  __shared__ int s_tile;
  const int tid = threadIdx.x;
  for (;;) {
    if (tid == 0)
      s_tile = (int)__hip_atomic_fetch_add(head, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const int it = s_tile;
    __syncthreads();
    if (it >= ntiles) break;
    const RTile t = tiles[it];
    const RClass &c = classes[t.cls];
    const int st = t.stage;
    const int L = c.L[st];
    const int gcount = t.count;
    if (st == 1 && wait) {
      if (tid == 0) {
        unsigned spins = 0;
        while (__hip_atomic_load(&deps[t.dep], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <
               need[t.dep]) {
          __builtin_amdgcn_s_sleep(2);
          if (++spins >= kSpinLimit) {
            __hip_atomic_store(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
          }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // see stage A's stores
      }
      __syncthreads();
    }
    HbmSide hin{}, hout{};
    int gin = 0, gout = 0;
    // load / store policy as in the uniform plans: ordinary loads where the first pass reads
    // short runs (single-stage tiles, stage B of the intermediate), ordinary stores for the
    // intermediate, non-temporal accesses otherwise
    if (c.nstage == 1) {  // G transforms of this size, listed by offset
      hin = HbmSide{const_cast<double2 *>(in), offs + t.base, 0, 1, nullptr, 0, 0, 1};
      hout = HbmSide{out, offs + t.base, 0, 1, nullptr, 0, 0};
    } else if (st == 0) {  // columns c of x[c + m Q] -> intermediate at c*P + k in out
      hin = HbmSide{const_cast<double2 *>(in) + t.base + t.sub0, nullptr, 1, c.Q, nullptr, 0, 0};
      hout = HbmSide{out + t.base + t.sub0 * c.P, nullptr, c.P, 1, nullptr, 0, 0, 1};
      gin = 1;
    } else {  // residue classes k1 + P i, in place in out
      hin = HbmSide{out + t.base + t.sub0, nullptr, 1, c.P, nullptr, 0, 0, 1};
      hout = HbmSide{out + t.base + t.sub0, nullptr, 1, c.P, nullptr, 0, 0};
      gin = gout = 1;
    }
    // one pass body per radix with its HBM / LDS sides chosen at run time: 50 KiB of code and
    // 117 VGPRs (four waves per SIMD) against 133 KiB and 129 (three) with a copy per side
    // combination; MI355X 1.51 -> 1.58 TB/s on the 10,000-transform batch
    run_stage_compact<FWD, RSET_MID, LDS_ELEMS / WG_THREADS>(lds, hin, hout, tid, WG_THREADS,
                                                            c.G[st], gcount, L, c.npass[st],
                                                            c.pass[st], gin, gout, tw, t.sub0,
                                                            st == 1 ? 1 : 0);
    if (c.nstage == 2 && st == 0 && wait) {
      __syncthreads();  // every thread's stores of the tile are issued
      if (tid == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");  // publish them
        __hip_atomic_fetch_add(&deps[t.dep], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    __syncthreads();
  }
}

__global__ void ragged_reset(unsigned *counters, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    counters[i] = 0u;
}

}  // namespace mfft

// ------------------------------------------------------------------------------ host side

struct mfft_ragged {
  int sign = -1;
  int64_t count = 0, total = 0;
  int ntiles = 0, ntwo = 0;
  int n1 = 0, n2 = 0;  // split mode: stage-A and single-stage tiles, then stage-B tiles
  bool split = true;
  int grid = 0;
  size_t lds_bytes = 0;
  mfft::RClass *d_classes = nullptr;
  mfft::RTile *d_tiles = nullptr;   // one-launch order
  mfft::RTile *d_tiles1 = nullptr;  // split mode, first launch
  mfft::RTile *d_tiles2 = nullptr;  // split mode, second launch
  int64_t *d_offs = nullptr;
  unsigned *d_counters = nullptr;
  unsigned *d_need = nullptr;
  double2 *d_tw = nullptr;
  std::vector<mfft_plan *> plans;  // one per class (twiddle layout source)
  std::map<int64_t, std::pair<std::vector<int>, int>> class_plan;  // size -> (radices, split)
};

namespace {

void free_ragged(mfft_ragged *r) {
  if (!r) return;
  (void)hipFree(r->d_classes);
  (void)hipFree(r->d_tiles);
  (void)hipFree(r->d_tiles1);
  (void)hipFree(r->d_tiles2);
  (void)hipFree(r->d_offs);
  (void)hipFree(r->d_counters);
  (void)hipFree(r->d_need);
  (void)hipFree(r->d_tw);
  for (auto *p : r->plans) mfft_destroy_plan(p);
  delete r;
}

}  // namespace

extern "C" {

mfft_ragged *mfft_plan_ragged(int64_t count, const int64_t *sizes, int sign, unsigned flags) {
  if (count < 0 || (count > 0 && !sizes) || (sign != -1 && sign != 1)) return nullptr;
  auto *r = new mfft_ragged;
  r->sign = sign;
  r->count = count;
  std::vector<int64_t> offs(count);
  std::map<int64_t, std::vector<int64_t>> by_size;  // size -> transform indices
  int64_t off = 0;
  for (int64_t i = 0; i < count; ++i) {
    offs[i] = off;
    off += sizes[i];
    by_size[sizes[i]].push_back(i);
  }
  r->total = off;

  std::vector<mfft::RClass> classes;
  std::vector<double2> tw;
  std::map<int64_t, int> cls_of;
  for (auto &kv : by_size) {
    mfft_plan *p = plan_many_internal(kv.first, sign, 1, 1, kv.first, 1, kv.first, flags, false);
    if (!p) {
      free_ragged(r);
      return nullptr;
    }
    r->plans.push_back(p);
    {
      int buf[64], st[4];
      const int nr = mfft_plan_radices(p, buf, 64);
      const int ns = mfft_plan_stages(p, st, 4);
      r->class_plan[kv.first] = {std::vector<int>(buf, buf + nr), ns == 2 ? st[0] : 0};
    }
    mfft::RClass c{};
    if (mfft::export_class(p, &c, &tw) != 0) {
      free_ragged(r);
      return nullptr;
    }
    cls_of[kv.first] = (int)classes.size();
    classes.push_back(c);
  }

  // tiles, largest class first
  std::vector<mfft::RTile> single, two;
  std::vector<int64_t> tile_offs;
  std::vector<unsigned> need;
  struct TwoT {
    std::vector<mfft::RTile> a, b;
  };
  std::vector<TwoT> twos;
  for (auto it = by_size.rbegin(); it != by_size.rend(); ++it) {
    const int ci = cls_of[it->first];
    const mfft::RClass &c = classes[ci];
    const auto &ids = it->second;
    if (c.nstage == 1) {
      const int G = c.G[0];
      for (size_t k = 0; k < ids.size(); k += G) {
        mfft::RTile t{};
        t.cls = ci;
        t.stage = 0;
        t.count = (int)std::min<size_t>(G, ids.size() - k);
        t.dep = -1;
        t.base = (int64_t)tile_offs.size();
        t.sub0 = 0;
        for (size_t q = k; q < k + t.count; ++q) tile_offs.push_back(offs[ids[q]]);
        single.push_back(t);
      }
    } else {
      for (int64_t id : ids) {
        TwoT tt;
        const int dep = (int)need.size();
        unsigned na = 0;
        for (int64_t c0 = 0; c0 < c.Q; c0 += c.G[0]) {
          mfft::RTile t{};
          t.cls = ci;
          t.stage = 0;
          t.count = (int)std::min<int64_t>(c.G[0], c.Q - c0);
          t.dep = dep;
          t.base = offs[id];
          t.sub0 = c0;
          tt.a.push_back(t);
          ++na;
        }
        for (int64_t k0 = 0; k0 < c.P; k0 += c.G[1]) {
          mfft::RTile t{};
          t.cls = ci;
          t.stage = 1;
          t.count = (int)std::min<int64_t>(c.G[1], c.P - k0);
          t.dep = dep;
          t.base = offs[id];
          t.sub0 = k0;
          tt.b.push_back(t);
        }
        need.push_back(na);
        twos.push_back(std::move(tt));
      }
    }
  }
  // Interleave: the stage-B tiles of a transform are queued only after at least LAG
  // further stage-A tiles, LAG a few times the number of resident workgroups, so that its
  // stage-A tiles have almost always finished when a workgroup takes its stage-B tiles.
  int dev0 = 0, ncu0 = 0;
  (void)hipGetDevice(&dev0);
  (void)hipDeviceGetAttribute(&ncu0, hipDeviceAttributeMultiprocessorCount, dev0);
  const size_t LAG = (size_t)std::max(ncu0, 1) * 2 * 3;
  std::vector<mfft::RTile> order;
  size_t pending = 0, a_since = 0;  // next transform whose B tiles are pending
  for (size_t i = 0; i < twos.size(); ++i) {
    order.insert(order.end(), twos[i].a.begin(), twos[i].a.end());
    a_since += twos[i].a.size();
    // release B tiles of earlier transforms once LAG A tiles were queued after them
    while (pending < i) {
      size_t after = 0;
      for (size_t q = pending + 1; q <= i; ++q) after += twos[q].a.size();
      if (after < LAG) break;
      order.insert(order.end(), twos[pending].b.begin(), twos[pending].b.end());
      ++pending;
    }
  }
  // the single-stage tiles never wait: queue them before the remaining B tiles
  order.insert(order.end(), single.begin(), single.end());
  for (; pending < twos.size(); ++pending)
    order.insert(order.end(), twos[pending].b.begin(), twos[pending].b.end());
  (void)a_since;
  r->ntiles = (int)order.size();
  r->ntwo = (int)need.size();
  // split mode: all stage-A tiles (largest first), then the single-stage tiles; then, in a
  // second launch, all stage-B tiles
  std::vector<mfft::RTile> first, second;
  for (auto &tt : twos) first.insert(first.end(), tt.a.begin(), tt.a.end());
  first.insert(first.end(), single.begin(), single.end());
  for (auto &tt : twos) second.insert(second.end(), tt.b.begin(), tt.b.end());
  r->n1 = (int)first.size();
  r->n2 = (int)second.size();
  {
    const char *e = std::getenv("MFFT_RAGGED_MODE");
    r->split = !(e && std::string(e) == "one");
  }

  auto up = [](void **dst, const void *src, size_t bytes) {
    if (bytes == 0) bytes = 16;
    if (hipMalloc(dst, bytes) != hipSuccess) return false;
    return src == nullptr || hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess;
  };
  bool ok = up((void **)&r->d_classes, classes.data(), classes.size() * sizeof(mfft::RClass)) &&
            up((void **)&r->d_tiles, order.data(), order.size() * sizeof(mfft::RTile)) &&
            up((void **)&r->d_tiles1, first.data(), first.size() * sizeof(mfft::RTile)) &&
            up((void **)&r->d_tiles2, second.data(), second.size() * sizeof(mfft::RTile)) &&
            up((void **)&r->d_offs, tile_offs.empty() ? nullptr : tile_offs.data(),
               tile_offs.size() * sizeof(int64_t)) &&
            up((void **)&r->d_need, need.empty() ? nullptr : need.data(),
               need.size() * sizeof(unsigned)) &&
            up((void **)&r->d_counters, nullptr, (3 + need.size()) * sizeof(unsigned)) &&
            up((void **)&r->d_tw, tw.empty() ? nullptr : tw.data(), tw.size() * sizeof(double2));
  if (!ok) {
    free_ragged(r);
    return nullptr;
  }
  for (auto *p : r->plans) mfft_destroy_plan(p);
  r->plans.clear();
  int dev = 0, ncu = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  size_t need_lds = 0;
  for (const auto &c : classes)
    for (int s2 = 0; s2 < c.nstage; ++s2)
      need_lds = std::max(need_lds, (size_t)mfft::stage_lds_elems(c.G[s2], c.L[s2], c.npass[s2]) *
                                        sizeof(double2));
  r->lds_bytes = need_lds;
  for (const void *k : {(const void *)mfft::ragged_kernel<true>,
                        (const void *)mfft::ragged_kernel<false>})
    (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, mfft::kLdsBytes);
  // a few resident workgroups per CU (about 35 KiB of LDS each); workgroups that do not fit
  // start when others finish and take what is left of the queue
  r->grid = std::max(1, std::min(r->ntiles, std::max(ncu, 1) * 4));
  return r;
}

int mfft_execute_ragged(const mfft_ragged *r, const void *in, void *out, void *stream) {
  if (!r || !in || !out) return MFFT_ERR_INVALID;
  if (in == out) return MFFT_ERR_UNSUPPORTED;  // stage A writes its intermediate into out
  if (r->ntiles == 0) return MFFT_SUCCESS;
  hipStream_t st = (hipStream_t)stream;
  const int nc = 3 + r->ntwo;  // [0] head, [1] error, [2] second head, [3 + t] stage-A done
  unsigned *ctr = r->d_counters;
  hipLaunchKernelGGL(mfft::ragged_reset, dim3((nc + 255) / 256), dim3(256), 0, st, ctr, nc);
  const double2 *x = (const double2 *)in;
  double2 *y = (double2 *)out;
  auto launch = [&](const mfft::RTile *tiles, int n, unsigned *head, int wait) {
    if (n == 0) return;
    const dim3 grid(std::max(1, std::min(n, r->grid))), block(mfft::WG_THREADS);
    if (r->sign < 0)
      hipLaunchKernelGGL(mfft::ragged_kernel<true>, grid, block, r->lds_bytes, st, x, y, r->d_tw,
                         r->d_classes, tiles, n, r->d_offs, head, ctr + 1, ctr + 3, r->d_need,
                         wait);
    else
      hipLaunchKernelGGL(mfft::ragged_kernel<false>, grid, block, r->lds_bytes, st, x, y, r->d_tw,
                         r->d_classes, tiles, n, r->d_offs, head, ctr + 1, ctr + 3, r->d_need,
                         wait);
  };
  if (r->split) {
    launch(r->d_tiles1, r->n1, ctr, 0);
    launch(r->d_tiles2, r->n2, ctr + 2, 0);
  } else {
    launch(r->d_tiles, r->ntiles, ctr, 1);
  }
  return hipGetLastError() == hipSuccess ? MFFT_SUCCESS : MFFT_ERR_HIP;
}

// Error flag of the last launch (1 if a stage-B wait hit its bound); synchronises.
int mfft_ragged_status(const mfft_ragged *r, void *stream) {
  if (!r) return MFFT_ERR_INVALID;
  unsigned flag = 0;
  if (hipStreamSynchronize((hipStream_t)stream) != hipSuccess ||
      hipMemcpy(&flag, r->d_counters + 1, sizeof(unsigned), hipMemcpyDeviceToHost) != hipSuccess)
    return MFFT_ERR_HIP;
  return flag ? MFFT_ERR_HIP : MFFT_SUCCESS;
}

int mfft_ragged_class_plan(const mfft_ragged *r, int64_t n, int *radices, int max, int *split) {
  if (!r) return -1;
  auto it = r->class_plan.find(n);
  if (it == r->class_plan.end()) return -1;
  const auto &rs = it->second.first;
  for (int i = 0; i < (int)rs.size() && i < max; ++i) radices[i] = rs[i];
  if (split) *split = it->second.second;
  return (int)rs.size();
}

void mfft_destroy_ragged(mfft_ragged *r) { free_ragged(r); }

}  // extern "C"
