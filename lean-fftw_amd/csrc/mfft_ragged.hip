// Ragged batches: thousands of transforms of different 13-smooth sizes, packed back to
// back, run by persistent launches.
//
// The host groups the transforms by size ("classes"; one plan per class, each plan the
// same Stockham schedule the uniform path uses) and cuts the work into tiles of at most
// 8 * RAGGED_THREADS = 4096 elements:
//   * a single-stage class (N <= LDS_ELEMS): a tile is G = 4096 / N transforms of that size,
//     listed by their element offsets;
//   * a two-stage class (N = P * Q): stage-A tiles of G_A = 4096 / P columns, stage-B tiles
//     of G_B = 4096 / Q residue classes; stage A writes its intermediate into the output
//     buffer and stage B transforms it there in place.
// Two queues, each largest class first (longest-processing-time order): the first holds
// every stage-A tile and then every single-stage tile, the second every stage-B tile. One
// persistent launch drains each queue (workgroups take tiles with an agent-scope atomic
// counter until it runs out); the kernel boundary between the two launches orders every
// stage-B tile after all stage-A tiles, so no workgroup ever waits for another. A HIP graph
// replays the counter reset and both launches as one unit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <string>
#include <vector>

#include "../../include/mfft.h"
#include "mfft_pass.h"

namespace mfft {

// Elements per thread of a ragged tile: a workgroup of NT threads holds G * L <= 8 NT points.
constexpr int RAGGED_EPT = 8;
// Workgroup size of both launches. MI355X, the 10,000-transform batch: 256 threads (tiles of
// 2048 points) 1.44-1.49 ms, 512 1.26-1.31 ms, 1024 1.37 ms; the longer rows of the wider
// tiles leave fewer partly used 128-byte lines (HBM traffic 7.4 -> 6.5 GB at 512).
constexpr int RAGGED_THREADS = 512;
// Work queues per launch. One queue shared by every workgroup: with eight (one per XCD, each
// size class's tiles in one queue, workgroup w draining queue w mod 8 first and then the
// others) the HBM traffic of the 10,000-transform batch fell from 6.47 to 5.13 GB but both
// launches ran about 40% longer on MI355X (1.26 -> 1.71 ms).
constexpr int RAGGED_QUEUES = 1;

struct RTile {
  int32_t cls;
  int32_t stage;  // 0: single stage or stage A, 1: stage B
  int32_t count;  // sub-problems (transforms, columns or residues) in the tile
  int32_t pad;
  int64_t base;   // single stage: index into offs[]; two stage: element offset of the transform
  int64_t sub0;   // two stage: first column / residue
};

// The queues of one launch: queue q holds tiles[off[q] .. off[q + 1]).
struct RQueues {
  int off[RAGGED_QUEUES + 1];
};

// Workgroup w drains queue w mod RAGGED_QUEUES first, then takes what is left of the others
// in turn. Only thread 0 touches the queue heads; the tile index reaches the workgroup
// through LDS.
template <bool FWD, int NT>
__global__ __launch_bounds__(NT) void ragged_kernel(
    const double2 *in, double2 *out, const double2 *__restrict__ tw,
    const RClass *__restrict__ classes, const RTile *__restrict__ tiles, const RQueues qs,
    const int64_t *__restrict__ offs, unsigned *heads) {
  extern __shared__ double2 lds[];
  __shared__ int s_tile;
  const int tid = threadIdx.x;
  int qi = 0;  // thread 0: queues before the (blockIdx.x + qi)-th are known to be empty
  for (;;) {
    if (tid == 0) {
      int found = -1;
      for (; qi < RAGGED_QUEUES && found < 0; ++qi) {
        const int q = (int)((blockIdx.x + qi) % RAGGED_QUEUES);
        const unsigned k =
            __hip_atomic_fetch_add(heads + q, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (k < (unsigned)(qs.off[q + 1] - qs.off[q])) {
          found = qs.off[q] + (int)k;
          break;  // this queue may hold more: try it again next time
        }
      }
      s_tile = found;
    }
    __syncthreads();
    const int it = s_tile;
    __syncthreads();
    if (it < 0) break;
    const RTile t = tiles[it];
    const RClass &c = classes[t.cls];
    const int st = t.stage;
    const int L = c.L[st];
    const int gcount = t.count;
    HbmSide hin{}, hout{};
    int gin = 0, gout = 0;
    // loads / stores: ordinary loads everywhere (stage A's strided rows included: MI355X
    // 1.267 -> 1.237 ms per 10,000-transform batch against non-temporal ones), ordinary
    // stores for the intermediate, non-temporal stores of the outputs
    if (c.nstage == 1) {  // G transforms of this size, listed by offset
      hin = HbmSide{const_cast<double2 *>(in), offs + t.base, 0, 1, nullptr, 0, 0, 1};
      hout = HbmSide{out, offs + t.base, 0, 1, nullptr, 0, 0};
    } else if (st == 0) {  // columns c of x[c + m Q] -> intermediate at c*P + k in out
      hin = HbmSide{const_cast<double2 *>(in) + t.base + t.sub0, nullptr, 1, c.Q, nullptr, 0, 0, 1};
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
    run_stage_compact<FWD, RSET_MID, RAGGED_EPT>(lds, hin, hout, tid, NT,
                                                            c.G[st], gcount, L, c.npass[st],
                                                            c.pass[st], gin, gout, tw, t.sub0,
                                                            st == 1 ? 1 : 0);
    __syncthreads();
  }
}

const void *ragged_kernel_for(bool fwd) {
  return fwd ? (const void *)ragged_kernel<true, RAGGED_THREADS>
             : (const void *)ragged_kernel<false, RAGGED_THREADS>;
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
  int n1 = 0, n2 = 0;  // tiles of the first launch (stage A, single stage) and of the second
  mfft::RQueues q1{}, q2{};  // their queues
  int grid = 0;
  size_t lds_bytes = 0;
  mfft::RClass *d_classes = nullptr;
  mfft::RTile *d_tiles1 = nullptr;
  mfft::RTile *d_tiles2 = nullptr;
  int64_t *d_offs = nullptr;
  unsigned *d_counters = nullptr;  // the queue heads: the first launch's, then the second's
  double2 *d_tw = nullptr;
  std::vector<mfft_plan *> plans;  // one per class (twiddle layout source)
  std::map<int64_t, std::pair<std::vector<int>, int>> class_plan;  // size -> (radices, split)
};

namespace mfft {

// What the queue builder needs of a size class.
struct RShape {
  int64_t n;
  int nstage;
  int64_t P, Q;  // two stages: N = P * Q
  int G0, G1;    // sub-problems per tile of each stage
};

// Longest-processing-time assignment of weighted items to nq queues: heaviest first (the
// items arrive largest class first, and the sort is stable), each to the least loaded queue
// (the lowest index on ties).
static std::vector<int> lpt_queues(const std::vector<int64_t> &weight, int nq) {
  std::vector<int> order(weight.size()), q(weight.size(), 0);
  for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return weight[a] > weight[b]; });
  std::vector<int64_t> load(nq, 0);
  for (int i : order) {
    const int best = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    q[i] = best;
    load[best] += weight[i];
  }
  return q;
}

// The work queues of a ragged batch's two launches. `cls` lists the distinct sizes in
// ascending order (class i has size cls[i].n). Transform i starts at element
// sum_{j<i} sizes[j]. Each launch has nq queues (one per XCD), and every class's tiles of a
// launch go to one queue, chosen by longest-processing-time on the class's points (so that
// neighbouring tiles of a transform, which share partly used 128-byte lines, and the
// transforms of a class, which share a twiddle table, meet in one XCD's L2). First launch,
// queue q: the stage-A tiles of its two-stage classes, then its single-stage tiles, each
// largest class first, a transform's tiles in sub-problem order; second launch: the stage-B
// tiles, largest class first. `first` and `second` hold the queues back to back, queue q at
// [qfirst[q], qfirst[q + 1]). Single-stage tiles list their transforms' offsets in `offs`.
// Returns false if a size has no class.
bool ragged_queues(int64_t count, const int64_t *sizes, const std::vector<RShape> &cls, int nq,
                   std::vector<RTile> &first, std::vector<RTile> &second,
                   std::vector<int> &qfirst, std::vector<int> &qsecond,
                   std::vector<int64_t> &offs) {
  if (nq < 1) return false;
  std::map<int64_t, std::vector<int64_t>> by_size;  // size -> transform offsets
  std::map<int64_t, int> cls_of;
  for (size_t i = 0; i < cls.size(); ++i) cls_of[cls[i].n] = (int)i;
  int64_t off = 0;
  for (int64_t i = 0; i < count; ++i) {
    if (!cls_of.count(sizes[i])) return false;
    by_size[sizes[i]].push_back(off);
    off += sizes[i];
  }
  offs.clear();
  // the tiles of every class present, largest class first
  struct Work {
    bool two;
    int64_t points;
    std::vector<RTile> a, b;  // stage A (or single stage), stage B
  };
  std::vector<Work> work;
  for (auto it = by_size.rbegin(); it != by_size.rend(); ++it) {
    const int ci = cls_of[it->first];
    const RShape &c = cls[ci];
    const auto &starts = it->second;
    Work w{c.nstage == 2, c.n * (int64_t)starts.size(), {}, {}};
    if (c.nstage == 1) {
      for (size_t k = 0; k < starts.size(); k += (size_t)c.G0) {
        RTile t{};
        t.cls = ci;
        t.stage = 0;
        t.count = (int)std::min<size_t>((size_t)c.G0, starts.size() - k);
        t.base = (int64_t)offs.size();
        offs.insert(offs.end(), starts.begin() + k, starts.begin() + k + t.count);
        w.a.push_back(t);
      }
    } else {
      for (int64_t start : starts) {
        for (int64_t c0 = 0; c0 < c.Q; c0 += c.G0) {
          RTile t{};
          t.cls = ci;
          t.stage = 0;
          t.count = (int)std::min<int64_t>(c.G0, c.Q - c0);
          t.base = start;
          t.sub0 = c0;
          w.a.push_back(t);
        }
        for (int64_t k0 = 0; k0 < c.P; k0 += c.G1) {
          RTile t{};
          t.cls = ci;
          t.stage = 1;
          t.count = (int)std::min<int64_t>(c.G1, c.P - k0);
          t.base = start;
          t.sub0 = k0;
          w.b.push_back(t);
        }
      }
    }
    work.push_back(std::move(w));
  }
  std::vector<int64_t> wa(work.size()), wb;
  std::vector<int> two_idx;
  for (size_t i = 0; i < work.size(); ++i) {
    wa[i] = work[i].points;
    if (work[i].two) {
      two_idx.push_back((int)i);
      wb.push_back(work[i].points);
    }
  }
  const std::vector<int> qa = lpt_queues(wa, nq), qb = lpt_queues(wb, nq);
  first.clear();
  second.clear();
  qfirst.assign(nq + 1, 0);
  qsecond.assign(nq + 1, 0);
  for (int q = 0; q < nq; ++q) {
    qfirst[q] = (int)first.size();
    for (int pass = 0; pass < 2; ++pass)  // two-stage classes first, then single-stage ones
      for (size_t i = 0; i < work.size(); ++i)
        if (qa[i] == q && work[i].two == (pass == 0))
          first.insert(first.end(), work[i].a.begin(), work[i].a.end());
    qsecond[q] = (int)second.size();
    for (size_t j = 0; j < two_idx.size(); ++j)
      if (qb[j] == q) second.insert(second.end(), work[two_idx[j]].b.begin(), work[two_idx[j]].b.end());
  }
  qfirst[nq] = (int)first.size();
  qsecond[nq] = (int)second.size();
  return true;
}

}  // namespace mfft

namespace {

void free_ragged(mfft_ragged *r) {
  if (!r) return;
  (void)hipFree(r->d_classes);
  (void)hipFree(r->d_tiles1);
  (void)hipFree(r->d_tiles2);
  (void)hipFree(r->d_offs);
  (void)hipFree(r->d_counters);
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
  std::map<int64_t, int> distinct;
  for (int64_t i = 0; i < count; ++i) {
    if (sizes[i] < 1) {
      free_ragged(r);
      return nullptr;
    }
    distinct[sizes[i]] = 0;
    r->total += sizes[i];
  }

  const int tile = mfft::RAGGED_EPT * mfft::RAGGED_THREADS;  // points per tile

  std::vector<mfft::RClass> classes;
  std::vector<mfft::RShape> shapes;
  std::vector<double2> tw;
  for (auto &kv : distinct) {
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
    // sub-problems per tile: as many as the tile holds (at most the stage's count)
    for (int s2 = 0; s2 < c.nstage; ++s2) {
      int64_t g = std::max(1, tile / c.L[s2]);
      if (c.nstage == 2) g = std::min<int64_t>(g, s2 == 0 ? c.Q : c.P);
      while (g > 1 && (int64_t)mfft::stage_lds_elems((int)g, c.L[s2], c.npass[s2]) *
                              (int64_t)sizeof(double2) > mfft::kLdsBytesMax)
        --g;
      c.G[s2] = (int)g;
      for (int q = 0; q < c.npass[s2]; ++q) c.pass[s2][q].m_G = mfft::magic32(c.G[s2]);
    }
    classes.push_back(c);
    shapes.push_back({kv.first, c.nstage, c.P, c.Q, c.G[0], c.nstage == 2 ? c.G[1] : 0});
  }
  std::vector<mfft::RTile> first, second;
  std::vector<int64_t> tile_offs;
  std::vector<int> qf, qs;
  if (!mfft::ragged_queues(count, sizes, shapes, mfft::RAGGED_QUEUES, first, second, qf, qs,
                           tile_offs)) {
    free_ragged(r);
    return nullptr;
  }
  r->n1 = (int)first.size();
  r->n2 = (int)second.size();
  for (int q = 0; q <= mfft::RAGGED_QUEUES; ++q) {
    r->q1.off[q] = qf[q];
    r->q2.off[q] = qs[q];
  }

  auto up = [](void **dst, const void *src, size_t bytes) {
    if (bytes == 0) bytes = 16;
    if (hipMalloc(dst, bytes) != hipSuccess) return false;
    return src == nullptr || hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess;
  };
  bool ok = up((void **)&r->d_classes, classes.data(), classes.size() * sizeof(mfft::RClass)) &&
            up((void **)&r->d_tiles1, first.data(), first.size() * sizeof(mfft::RTile)) &&
            up((void **)&r->d_tiles2, second.data(), second.size() * sizeof(mfft::RTile)) &&
            up((void **)&r->d_offs, tile_offs.empty() ? nullptr : tile_offs.data(),
               tile_offs.size() * sizeof(int64_t)) &&
            up((void **)&r->d_counters, nullptr, 2 * mfft::RAGGED_QUEUES * sizeof(unsigned)) &&
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
  for (bool fwd : {true, false})
    (void)hipFuncSetAttribute(mfft::ragged_kernel_for(fwd),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)need_lds);
  // two resident workgroups per CU (tiles of up to 70 KiB of LDS); workgroups that do not
  // fit start when others finish and take what is left of the queue
  r->grid = std::max(1, std::max(ncu, 1) * (1024 / mfft::RAGGED_THREADS));
  return r;
}

int mfft_execute_ragged(const mfft_ragged *r, const void *in, void *out, void *stream) {
  if (!r || !in || !out) return MFFT_ERR_INVALID;
  if (in == out) return MFFT_ERR_UNSUPPORTED;  // stage A writes its intermediate into out
  if (r->n1 + r->n2 == 0) return MFFT_SUCCESS;
  hipStream_t st = (hipStream_t)stream;
  unsigned *ctr = r->d_counters;  // the first launch's queue heads, then the second's
  {
    int nc = 2 * mfft::RAGGED_QUEUES;
    void *args[] = {(void *)&ctr, (void *)&nc};
    if (hipLaunchKernel((const void *)mfft::ragged_reset, dim3(1), dim3(256), args, 0, st) !=
        hipSuccess)
      return MFFT_ERR_HIP;
  }
  const double2 *x = (const double2 *)in;
  double2 *y = (double2 *)out;
  const void *k = mfft::ragged_kernel_for(r->sign < 0);
  auto launch = [&](const mfft::RTile *tiles, int n, const mfft::RQueues &qs, unsigned *heads) {
    if (n == 0) return hipSuccess;
    void *args[] = {(void *)&x, (void *)&y, (void *)&r->d_tw, (void *)&r->d_classes,
                    (void *)&tiles, (void *)&qs, (void *)&r->d_offs, (void *)&heads};
    return hipLaunchKernel(k, dim3(std::max(1, std::min(n, r->grid))), dim3(mfft::RAGGED_THREADS),
                           args, r->lds_bytes, st);
  };
  if (launch(r->d_tiles1, r->n1, r->q1, ctr) != hipSuccess ||
      launch(r->d_tiles2, r->n2, r->q2, ctr + mfft::RAGGED_QUEUES) != hipSuccess)
    return MFFT_ERR_HIP;
  return MFFT_SUCCESS;
}

// Synchronise the stream of the last launches and report whether it completed.
int mfft_ragged_status(const mfft_ragged *r, void *stream) {
  if (!r) return MFFT_ERR_INVALID;
  return hipStreamSynchronize((hipStream_t)stream) == hipSuccess ? MFFT_SUCCESS : MFFT_ERR_HIP;
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

int64_t mfft_ragged_queue(int64_t count, const int64_t *sizes, int64_t nclass,
                          const int64_t *shapes, int64_t nqueue, int64_t *tiles,
                          int64_t max_tiles, int64_t *n_first, int64_t *qoff, int64_t *offs,
                          int64_t max_offs) {
  if (count < 0 || nclass < 0 || (count && !sizes) || (nclass && !shapes) || nqueue < 1 ||
      nqueue > 64)
    return -1;
  std::vector<mfft::RShape> cls;
  for (int64_t i = 0; i < nclass; ++i) {
    const int64_t *s = shapes + 6 * i;
    if (s[1] < 1 || s[1] > 2 || s[4] < 1 || (s[1] == 2 && (s[5] < 1 || s[2] * s[3] != s[0])))
      return -1;
    if (i > 0 && s[0] <= shapes[6 * (i - 1)]) return -1;  // ascending distinct sizes
    cls.push_back({s[0], (int)s[1], s[2], s[3], (int)s[4], (int)s[5]});
  }
  std::vector<mfft::RTile> first, second;
  std::vector<int64_t> o;
  std::vector<int> qf, qs;
  if (!mfft::ragged_queues(count, sizes, cls, (int)nqueue, first, second, qf, qs, o)) return -1;
  if (n_first) *n_first = (int64_t)first.size();
  if (qoff)
    for (int64_t q = 0; q <= nqueue; ++q) {
      qoff[q] = qf[q];
      qoff[nqueue + 1 + q] = (int64_t)first.size() + qs[q];
    }
  first.insert(first.end(), second.begin(), second.end());
  for (size_t i = 0; i < first.size() && (int64_t)i < max_tiles && tiles; ++i) {
    const mfft::RTile &t = first[i];
    int64_t *d = tiles + 5 * i;
    d[0] = t.cls;
    d[1] = t.stage;
    d[2] = t.count;
    d[3] = t.base;
    d[4] = t.sub0;
  }
  for (size_t i = 0; i < o.size() && (int64_t)i < max_offs && offs; ++i) offs[i] = o[i];
  return (int64_t)first.size();
}

void mfft_destroy_ragged(mfft_ragged *r) { free_ragged(r); }

}  // extern "C"
