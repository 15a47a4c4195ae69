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
// Tiles are queued largest class first (longest-processing-time order). The stage-B tiles
// of a transform are queued a few transforms after its stage-A tiles, so they rarely wait,
// and its intermediate is still in the Infinity Cache when they run.
//
// Workgroups take tiles from the queue with an agent-scope atomic counter until it runs
// out. A stage-B tile waits (bounded spin) until the transform's stage-A counter, raised
// with release semantics after each stage-A tile, reaches the number of stage-A tiles;
// stage-A tiles are always queued before the stage-B tiles of their transform and never
// wait, so every wait ends. If a spin bound is ever hit the launch records an error flag
// and still exits.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <map>
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
    const int64_t *__restrict__ offs, unsigned *counters /* [0]: queue head, [1]: error,
    [2 + t]: stage-A tiles done of transform t */, const unsigned *__restrict__ need) {
  extern __shared__ double2 lds[];
  __shared__ int s_tile;
  const int tid = threadIdx.x;
  for (;;) {
    if (tid == 0)
      s_tile = (int)__hip_atomic_fetch_add(&counters[0], 1u, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const int it = s_tile;
    __syncthreads();
    if (it >= ntiles) break;
    const RTile t = tiles[it];
    const RClass &c = classes[t.cls];
    const int st = t.stage;
    const int L = c.L[st];
    const int gcount = t.count;
    if (st == 1) {
      if (tid == 0) {
        unsigned spins = 0;
        while (__hip_atomic_load(&counters[2 + t.dep], __ATOMIC_ACQUIRE,
                                 __HIP_MEMORY_SCOPE_AGENT) < need[t.dep]) {
          __builtin_amdgcn_s_sleep(2);
          if (++spins >= kSpinLimit) {
            __hip_atomic_store(&counters[1], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
          }
        }
      }
      __syncthreads();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // see stage A's stores
    }
    double2 *const buf0 = lds, *const buf1 = lds + LDS_ELEMS;
    HbmSide hin{}, hout{};
    int gin = 0, gout = 0;
    if (c.nstage == 1) {  // G transforms of this size, listed by offset
      hin = HbmSide{const_cast<double2 *>(in), offs + t.base, 0, 1, nullptr, 0, 0};
      hout = HbmSide{out, offs + t.base, 0, 1, nullptr, 0, 0};
    } else if (st == 0) {  // columns c of x[c + m Q] -> intermediate at c*P + k in out
      hin = HbmSide{const_cast<double2 *>(in) + t.base + t.sub0, nullptr, 1, c.Q, nullptr, 0, 0};
      hout = HbmSide{out + t.base + t.sub0 * c.P, nullptr, c.P, 1, nullptr, 0, 0};
      gin = 1;
    } else {  // residue classes k1 + P i, in place in out
      hin = HbmSide{out + t.base + t.sub0, nullptr, 1, c.P, nullptr, 0, 0};
      hout = HbmSide{out + t.base + t.sub0, nullptr, 1, c.P, nullptr, 0, 0};
      gin = gout = 1;
    }
    run_stage<FWD>(buf0, buf1, hin, hout, tid, WG_THREADS, c.G[st], gcount, L, c.npass[st],
                   c.pass[st], gin, gout, tw, t.sub0, st == 1 ? 1 : 0);
    if (c.nstage == 2 && st == 0) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");  // publish this tile's stores
      __syncthreads();
      if (tid == 0)
        __hip_atomic_fetch_add(&counters[2 + t.dep], 1u, __ATOMIC_RELEASE,
                               __HIP_MEMORY_SCOPE_AGENT);
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
  int grid = 0;
  size_t lds_bytes = 0;
  mfft::RClass *d_classes = nullptr;
  mfft::RTile *d_tiles = nullptr;
  int64_t *d_offs = nullptr;
  unsigned *d_counters = nullptr;
  unsigned *d_need = nullptr;
  double2 *d_tw = nullptr;
  std::vector<mfft_plan *> plans;  // one per class (twiddle layout source)
};

namespace {

void free_ragged(mfft_ragged *r) {
  if (!r) return;
  (void)hipFree(r->d_classes);
  (void)hipFree(r->d_tiles);
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
    mfft_plan *p = mfft_plan_dft_1d(kv.first, sign, 1, flags);
    if (!p) {
      free_ragged(r);
      return nullptr;
    }
    r->plans.push_back(p);
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
  // interleave: stage-B tiles of transform i follow the stage-A tiles of transform i+LAG
  const size_t LAG = 4;
  std::vector<mfft::RTile> order;
  for (size_t i = 0; i < twos.size(); ++i) {
    order.insert(order.end(), twos[i].a.begin(), twos[i].a.end());
    if (i >= LAG) order.insert(order.end(), twos[i - LAG].b.begin(), twos[i - LAG].b.end());
  }
  for (size_t i = twos.size() > LAG ? twos.size() - LAG : 0; i < twos.size(); ++i)
    order.insert(order.end(), twos[i].b.begin(), twos[i].b.end());
  order.insert(order.end(), single.begin(), single.end());
  r->ntiles = (int)order.size();
  r->ntwo = (int)need.size();

  auto up = [](void **dst, const void *src, size_t bytes) {
    if (bytes == 0) bytes = 16;
    if (hipMalloc(dst, bytes) != hipSuccess) return false;
    return src == nullptr || hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess;
  };
  bool ok = up((void **)&r->d_classes, classes.data(), classes.size() * sizeof(mfft::RClass)) &&
            up((void **)&r->d_tiles, order.data(), order.size() * sizeof(mfft::RTile)) &&
            up((void **)&r->d_offs, tile_offs.empty() ? nullptr : tile_offs.data(),
               tile_offs.size() * sizeof(int64_t)) &&
            up((void **)&r->d_need, need.empty() ? nullptr : need.data(),
               need.size() * sizeof(unsigned)) &&
            up((void **)&r->d_counters, nullptr, (2 + need.size()) * sizeof(unsigned)) &&
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
  r->lds_bytes = 2 * (size_t)mfft::LDS_ELEMS * sizeof(double2);
  // two resident workgroups per CU with 64 KiB of LDS each
  r->grid = std::max(1, std::min(r->ntiles, std::max(ncu, 1) * 2));
  return r;
}

int mfft_execute_ragged(const mfft_ragged *r, const void *in, void *out, void *stream) {
  if (!r || !in || !out) return MFFT_ERR_INVALID;
  if (in == out) return MFFT_ERR_UNSUPPORTED;  // stage A writes its intermediate into out
  if (r->ntiles == 0) return MFFT_SUCCESS;
  hipStream_t st = (hipStream_t)stream;
  const int nc = 2 + r->ntwo;
  hipLaunchKernelGGL(mfft::ragged_reset, dim3((nc + 255) / 256), dim3(256), 0, st, r->d_counters,
                     nc);
  if (r->sign < 0)
    hipLaunchKernelGGL(mfft::ragged_kernel<true>, dim3(r->grid), dim3(mfft::WG_THREADS),
                       r->lds_bytes, st, (const double2 *)in, (double2 *)out, r->d_tw,
                       r->d_classes, r->d_tiles, r->ntiles, r->d_offs, r->d_counters, r->d_need);
  else
    hipLaunchKernelGGL(mfft::ragged_kernel<false>, dim3(r->grid), dim3(mfft::WG_THREADS),
                       r->lds_bytes, st, (const double2 *)in, (double2 *)out, r->d_tw,
                       r->d_classes, r->d_tiles, r->ntiles, r->d_offs, r->d_counters, r->d_need);
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

void mfft_destroy_ragged(mfft_ragged *r) { free_ragged(r); }

}  // extern "C"
