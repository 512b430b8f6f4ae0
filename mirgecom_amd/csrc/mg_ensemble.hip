// Batched ensemble of independent cases: many uiuc_7sp mixture
// Navier-Stokes cases of uneven mesh size and polynomial order evaluated in
// persistent launches driven by per-case descriptors, instead of one launch
// sequence per case (the reference runs one driver per case:
// mirgecom/steppers.py:350-449 around mirgecom/navierstokes.py:350-575).
//
// One right-hand side of every case = five phases, each one launch over the
// work items of all cases (the element phases G and D, untuned, one launch
// per polynomial order):
//   T   Newton temperature from each case's seed   (items: case x 256 nodes)
//   G   FP32 gradients, FP32 contraction            (grad_body<!FLUX, float>)
//   N   FP32 nodal net flux, node by node           (node_flux_body<F32>)
//   D   divergence + face fluxes, FP32 contraction   (rhs_body<kFlux, float>)
//   S   species sources added into the RHS: the learned surrogate with its
//       gate (infer_body), or the exact rates
// The phase bodies are the same device functions the single-case entry
// points run, so each case's result equals its single-case evaluation bit
// for bit.  Workgroups are persistent (grid = workgroups per CU x CUs) and
// stride over the item list; items are ordered by decreasing order p (the
// most expensive first) and case by case within an order.
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "mgdg_host.h"
#include "mgchem_kernels.h"

namespace mgens {

// Descriptors are read from memory: their pointers are global-qualified
// (mgdg::KArgsT<1>), so that the bodies access memory with global_load /
// global_store as the single-case kernels do (flat accesses made the
// divergence phase 35-80 % slower).
using KArgs = mgdg::KArgsT<1>;
using mgdg::kBlock;

constexpr int kDim = 3, kNsp = 7, kMix = 1;
using R = float;

struct Item {
  int32_t icase;
  int32_t p;
  int64_t blk;
};

// A case's run of work items: items first .. first + (its block count) - 1
// of a launch belong to case icase (order p).  A workgroup finds its item's
// case by binary search over the launch's few segments (wave-uniform, scalar
// loads of a small table that stays in the caches), instead of loading a
// per-item record that would be a cache miss for every workgroup.
struct Seg {
  int32_t icase;
  int32_t p;
  int64_t first;
};

struct PointCase {
  __attribute__((address_space(1))) const double* q;
  __attribute__((address_space(1))) const double* tseed;
  __attribute__((address_space(1))) double* tout;
  int64_t n;
};
using InferArgs = mgchem::InferArgsT<1>;

// Item `it` of a launch (the same for the whole workgroup): the case
// descriptor is then read with scalar loads.
__device__ __forceinline__ Item find_item(const Seg* __restrict__ segs, int nseg,
                                          int64_t it) {
  int lo = 0, hi = nseg - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (segs[mid].first <= it)
      lo = mid;
    else
      hi = mid - 1;
  }
  lo = __builtin_amdgcn_readfirstlane(lo);
  const Seg sg = segs[lo];
  Item w;
  w.icase = __builtin_amdgcn_readfirstlane(sg.icase);
  w.p = __builtin_amdgcn_readfirstlane(sg.p);
  const uint64_t b = uint64_t(it - sg.first);
  w.blk = int64_t((uint64_t(uint32_t(__builtin_amdgcn_readfirstlane(uint32_t(b >> 32)))) << 32) |
                  uint32_t(__builtin_amdgcn_readfirstlane(uint32_t(b))));
  return w;
}

template <int P>
using GCfg = mgdg::GradCfg<kDim, P, kNsp, kMix, R>;
template <int P>
using DCfg = mgdg::RhsCfg<kDim, P, kNsp, kMix, mgdg::kFlux, R>;

template <int PH, int P>
__device__ __forceinline__ void run_body(const KArgs& a, int64_t blk) {
  if constexpr (PH == 0)
    mgdg::grad_body<kDim, P, kNsp, kMix, false, R>(a, (const double*)a.ops, (double*)a.out,
                                                   blk);
  else
    mgdg::rhs_body<kDim, P, kNsp, kMix, mgdg::kFlux, R, false>(a, (const double*)a.ops,
                                                               (double*)a.out, blk);
}

template <int PH, int PSEL>
__device__ __forceinline__ void run_item(const KArgs* __restrict__ cases, const Item& w) {
  const KArgs& a = cases[w.icase];
  if constexpr (PSEL == 0) {
    switch (w.p) {
      case 1: run_body<PH, 1>(a, w.blk); break;
      case 2: run_body<PH, 2>(a, w.blk); break;
      case 3: run_body<PH, 3>(a, w.blk); break;
      default: run_body<PH, 4>(a, w.blk); break;
    }
  } else {
    run_body<PH, PSEL>(a, w.blk);
  }
}

// PH 0: gradients, 1: divergence.  PSEL 0: any order (the
// body is selected per item), else only items of order PSEL.  PERSIST:
// workgroups stride over the items (grid = workgroups per CU x CUs; the
// loop keeps more values live, ~50 more VGPRs for the gradient phase), else
// one workgroup per item (the single-case kernels' register budget).
// Both schedules keep the items of one XCD contiguous (items are ordered
// so that consecutive ones are mesh neighbours; see mgdg::xcd_block): with
// one workgroup per item through the same bijection as the single-case
// kernels, persistently by giving XCD x the x-th eighth of the item list.
template <int PH, int PSEL, bool PERSIST>
__global__ __launch_bounds__(kBlock) void ens_dg_kernel(
    const KArgs* __restrict__ cases, const Seg* __restrict__ segs, int nseg, int64_t base,
    int64_t nitems) {
  if constexpr (PERSIST) {
    const int64_t G = gridDim.x, b = blockIdx.x;
    int64_t lo = b, hi = nitems, step = G;
    if (G >= mgdg::kXcds) {
      const int64_t x = b % mgdg::kXcds, i = b / mgdg::kXcds;
      step = G / mgdg::kXcds + (x < G % mgdg::kXcds ? 1 : 0);
      hi = nitems * (x + 1) / mgdg::kXcds;
      lo = nitems * x / mgdg::kXcds + i;
    }
    for (int64_t it = lo; it < hi; it += step) {
      run_item<PH, PSEL>(cases, find_item(segs, nseg, base + it));
      __syncthreads();   // the next item reuses the LDS
    }
  } else {
    const int64_t it = mgdg::xcd_block();
    if (it < nitems) run_item<PH, PSEL>(cases, find_item(segs, nseg, base + it));
  }
}

__global__ __launch_bounds__(kBlock) void ens_temperature_kernel(
    const PointCase* __restrict__ cases, const Seg* __restrict__ segs, int nseg,
    int64_t nitems, int niter) {
  for (int64_t it = blockIdx.x; it < nitems; it += gridDim.x) {
    const Item w = find_item(segs, nseg, it);
    const PointCase c = cases[w.icase];
    const int64_t g = w.blk * kBlock + threadIdx.x;
    if (g < c.n) {
      const int64_t n = c.n;
      const double rho = c.q[g];
      double mm = 0.0;
#pragma unroll
      for (int d = 0; d < kDim; ++d) mm += c.q[(2 + d) * n + g] * c.q[(2 + d) * n + g];
      const double e = (c.q[n + g] - 0.5 * mm / rho) / rho;
      double y[kNsp];
#pragma unroll
      for (int k = 0; k < kNsp; ++k) y[k] = c.q[(2 + kDim + k) * n + g] / rho;
      c.tout[g] = mgdg::mix_temperature(e, c.tseed[g], y, niter);
    }
  }
}

// Phase N: the FP32 nodal net flux of every node (items: case x 256 nodes;
// descriptors: the gradient phase's, which carry grad32 and fnet32_out).
__global__ __launch_bounds__(kBlock, 2) void ens_node_flux_kernel(
    const KArgs* __restrict__ cases, const Seg* __restrict__ segs, int nseg,
    int64_t nitems) {
  __shared__ double s_pref;
  for (int64_t it = blockIdx.x; it < nitems; it += gridDim.x) {
    const Item w = find_item(segs, nseg, it);
    const KArgs& c = cases[w.icase];
    __syncthreads();   // the previous item's s_pref has been read
    if (threadIdx.x == 0) s_pref = mgdg::reference_pressure<kDim, kNsp, kMix>(c);
    __syncthreads();
    const int64_t g = w.blk * kBlock + threadIdx.x;
    if (g < c.nnodes) mgdg::node_flux_body<kDim, kNsp, kMix, true>(c, g, s_pref);
  }
}

__global__ __launch_bounds__(kBlock) void ens_exact_source_kernel(
    const InferArgs* __restrict__ cases, const Seg* __restrict__ segs, int nseg,
    int64_t nitems) {
  for (int64_t it = blockIdx.x; it < nitems; it += gridDim.x) {
    const Item w = find_item(segs, nseg, it);
    const InferArgs& c = cases[w.icase];
    const int64_t g = w.blk * kBlock + threadIdx.x;
    if (g < c.n) {
      const int64_t n = c.n;
      const double rho = c.q[g];
      double y[kNsp], s[kNsp];
#pragma unroll
      for (int k = 0; k < kNsp; ++k) y[k] = c.q[(2 + kDim + k) * n + g] / rho;
      mgdg::mix_species_source(rho, c.temp[g], y, s);
#pragma unroll
      for (int k = 0; k < kNsp; ++k) c.out[(2 + kDim + k) * n + g] += s[k];
    }
  }
}

__global__ __launch_bounds__(256, 2) void ens_infer_kernel(
    const InferArgs* __restrict__ cases, const Seg* __restrict__ segs, int nseg,
    int64_t nitems) {
  for (int64_t it = blockIdx.x; it < nitems; it += gridDim.x) {
    const Item w = find_item(segs, nseg, it);
    mgchem::infer_body(cases[w.icase], w.blk);
    __syncthreads();
  }
}

}  // namespace mgens

namespace {

using mgens::Item;
using mgens::Seg;
constexpr int kBlock = mgdg::kBlock;

// Host image of mgens::PointCase (generic pointers, same layout)
struct PointHost {
  const double* q;
  const double* tseed;
  double* tout;
  int64_t n;
};
static_assert(sizeof(PointHost) == sizeof(mgens::PointCase), "point descriptor layout");

template <int P>
int grad_epb() { return mgens::GCfg<P>::EPB; }
template <int P>
int div_epb() { return mgens::DCfg<P>::EPB; }
int epb_of(int phase, int p) {
  switch (p) {
    case 1: return phase ? div_epb<1>() : grad_epb<1>();
    case 2: return phase ? div_epb<2>() : grad_epb<2>();
    case 3: return phase ? div_epb<3>() : grad_epb<3>();
    default: return phase ? div_epb<4>() : grad_epb<4>();
  }
}
template <int P>
size_t grad_lds() { return mgens::GCfg<P>::LDS; }
template <int P>
size_t div_lds() { return mgens::DCfg<P>::LDS; }
size_t lds_of(int phase, int p) {
  if (p == 0) {
    size_t m = 0;
    for (int q = 1; q <= 4; ++q) m = std::max(m, lds_of(phase, q));
    return m;
  }
  switch (p) {
    case 1: return phase ? div_lds<1>() : grad_lds<1>();
    case 2: return phase ? div_lds<2>() : grad_lds<2>();
    case 3: return phase ? div_lds<3>() : grad_lds<3>();
    default: return phase ? div_lds<4>() : grad_lds<4>();
  }
}

// Layout of the plan's device buffer.
struct Layout {
  int64_t grad_tab, div_tab, pt_tab, src_tab, grad_items, div_items, pt_items, src_items,
      bytes;
};

int64_t align256(int64_t x) { return (x + 255) / 256 * 256; }

// Work items of one case in a phase: element blocks (0: gradients, 1:
// divergence), node blocks (2: Newton), row blocks (3: sources; with the
// surrogate, whole elements of at most 128 nodes, else node blocks).
int64_t case_items(const mg_mesh& m, int phase, bool surrogate) {
  if (phase < 2) {
    const int epb = epb_of(phase, m.order);
    return (m.nelements + epb - 1) / epb;
  }
  const int64_t n = m.nelements * m.np;
  const int64_t per =
      (phase == 3 && surrogate) ? (mgchem::kRows / m.np) * m.np : int64_t(kBlock);
  return (n + per - 1) / per;
}

Layout layout(int32_t ncases, const mg_ens_case* cases) {
  Layout L{};
  int64_t off = 0;
  L.grad_tab = off;
  off = align256(off + int64_t(sizeof(mgdg::KArgs)) * ncases);
  L.div_tab = off;
  off = align256(off + int64_t(sizeof(mgdg::KArgs)) * ncases);
  L.pt_tab = off;
  off = align256(off + int64_t(sizeof(mgens::PointCase)) * ncases);
  L.src_tab = off;
  off = align256(off + int64_t(sizeof(mgchem::InferArgs)) * ncases);
  L.grad_items = off;
  off = align256(off + int64_t(sizeof(Seg)) * ncases);
  L.div_items = off;
  off = align256(off + int64_t(sizeof(Seg)) * ncases);
  L.pt_items = off;
  off = align256(off + int64_t(sizeof(Seg)) * ncases);
  L.src_items = off;
  off = align256(off + int64_t(sizeof(Seg)) * ncases);
  L.bytes = off;
  return L;
}

mg_status check_cases(int32_t ncases, const mg_ens_case* cases) {
  if (ncases < 1 || !cases) return fail(MG_ERR_ARG, "no cases");
  for (int c = 0; c < ncases; ++c) {
    const mg_ens_case& k = cases[c];
    mg_status s = check_mesh(&k.mesh);
    if (s) return fail(s, "case " + std::to_string(c) + ": " + mgdg_shared::err());
    if (k.mesh.dim != 3) return fail(MG_ERR_UNSUPPORTED, "ensemble cases are 3-D");
    if (k.mesh.nbnd_nodes != 0)
      return fail(MG_ERR_UNSUPPORTED, "ensemble cases are periodic (no boundary nodes)");
    if ((s = check_gas(&k.gas, true))) return s;
    if (k.gas.eos_kind != 1 || k.gas.transport_kind == 3)
      return fail(MG_ERR_UNSUPPORTED,
                  "ensemble cases: uiuc_7sp mixture with a non-mixture-averaged transport");
    if (!k.q || !k.t_seed || !k.temperature || !k.fnet || !k.grad || !k.rhs)
      return fail(MG_ERR_ARG, "case " + std::to_string(c) + ": NULL buffer");
  }
  return MG_OK;
}

mg_status dg_attrs() {
  static bool done = false;
  if (done) return MG_OK;
  mg_status s;
  // fused (any order) and per-order kernels of both phases, both schedules
#define MG_ENS_ATTR(PH, PS)                                                    \
  if ((s = set_lds(mgens::ens_dg_kernel<PH, PS, true>, lds_of(PH, PS)))) return s; \
  if ((s = set_lds(mgens::ens_dg_kernel<PH, PS, false>, lds_of(PH, PS)))) return s;
  MG_ENS_ATTR(0, 0) MG_ENS_ATTR(0, 1) MG_ENS_ATTR(0, 2) MG_ENS_ATTR(0, 3) MG_ENS_ATTR(0, 4)
  MG_ENS_ATTR(1, 0) MG_ENS_ATTR(1, 1) MG_ENS_ATTR(1, 2) MG_ENS_ATTR(1, 3) MG_ENS_ATTR(1, 4)
#undef MG_ENS_ATTR
  if ((s = set_lds(mgens::ens_infer_kernel, mgchem::kInferLds))) return s;
  done = true;
  return MG_OK;
}

mg_status launch_dg(int phase, int psel, bool persist, int grid, size_t lds,
                    const mgens::KArgs* tab, const Seg* segs, int nseg, int64_t base,
                    int64_t n, hipStream_t st) {
  if (n == 0) return MG_OK;
  if (!persist && n > 0x7fffffff) return fail(MG_ERR_ARG, "too many items for one launch");
#define MG_ENS(PH, PS)                                                                    \
  if (phase == PH && psel == PS) {                                                        \
    if (persist)                                                                          \
      hipLaunchKernelGGL((mgens::ens_dg_kernel<PH, PS, true>), dim3(unsigned(grid)),       \
                         dim3(kBlock), lds, st, tab, segs, nseg, base, n);                \
    else                                                                                  \
      hipLaunchKernelGGL((mgens::ens_dg_kernel<PH, PS, false>), dim3(unsigned(n)),         \
                         dim3(kBlock), lds, st, tab, segs, nseg, base, n);                \
    return check_launch("ens_dg_kernel");                                                 \
  }
  MG_ENS(0, 0) MG_ENS(0, 1) MG_ENS(0, 2) MG_ENS(0, 3) MG_ENS(0, 4)
  MG_ENS(1, 0) MG_ENS(1, 1) MG_ENS(1, 2) MG_ENS(1, 3) MG_ENS(1, 4)
#undef MG_ENS
  return fail(MG_ERR_ARG, "phase / order selector");
}

int device_cus() {
  static int cus = 0;
  if (cus == 0) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
      cus = 256;
  }
  return cus;
}

int grid_for(int64_t nitems, int wg_per_cu, int cus) {
  const int64_t g = int64_t(std::max(1, wg_per_cu)) * cus;
  return int(std::max<int64_t>(1, std::min<int64_t>(g, nitems)));
}

}  // namespace

extern "C" {

int64_t mg_ensemble_plan_bytes(int32_t ncases, const mg_ens_case* cases) {
  if (check_cases(ncases, cases)) return -1;
  return layout(ncases, cases).bytes;
}

mg_status mg_ensemble_prepare(int32_t ncases, const mg_ens_case* cases, int32_t niter,
                              double dt, const float* mlp_params, const uint16_t* mlp_wbf16,
                              const mg_chem_norm* norm, int32_t* fallback_count,
                              const mg_ens_tuning* tuning, void* dev, int64_t dev_bytes,
                              mg_ens_plan* plan, void* stream) {
  mg_status s;
  if ((s = check_cases(ncases, cases))) return s;
  if ((s = ensure_mech())) return s;
  if (!plan || !dev) return fail(MG_ERR_ARG, "plan / device buffer NULL");
  if (niter < 0) return fail(MG_ERR_ARG, "niter must be >= 0");
  const bool surrogate = mlp_params != nullptr;
  if (surrogate) {
    if (!mlp_wbf16 || !(dt > 0.0)) return fail(MG_ERR_ARG, "surrogate needs weights and dt > 0");
    if ((s = check_norm(norm))) return s;
  }
  const Layout L = layout(ncases, cases);
  if (dev_bytes < L.bytes) return fail(MG_ERR_ARG, "plan buffer too small");
  if ((s = dg_attrs())) return s;
  std::vector<char> host(size_t(L.bytes), 0);
  // host tables: the generic argument blocks, byte-identical to the device
  // side's global-qualified ones
  auto* gt = reinterpret_cast<mgdg::KArgs*>(host.data() + L.grad_tab);
  auto* dt_ = reinterpret_cast<mgdg::KArgs*>(host.data() + L.div_tab);
  auto* pt = reinterpret_cast<PointHost*>(host.data() + L.pt_tab);
  auto* st_ = reinterpret_cast<mgchem::InferArgs*>(host.data() + L.src_tab);
  for (int c = 0; c < ncases; ++c) {
    const mg_ens_case& k = cases[c];
    mgdg::KArgs a = make_args(&k.mesh, &k.gas);
    a.q = k.q;
    a.temp = k.temperature;
    a.flux_kind = MG_FLUX_RUSANOV;
    a.inviscid_on = 1;
    mgdg::KArgs g = a;   // gradient and node-flux phases
    g.grad32_out = k.grad;
    g.grad32 = k.grad;
    g.fnet32_out = k.fnet;
    g.out = nullptr;
    gt[c] = g;
    mgdg::KArgs d = a;
    d.fnet32 = k.fnet;
    d.out = k.rhs;
    dt_[c] = d;
    pt[c] = PointHost{k.q, k.t_seed, k.temperature, k.mesh.nelements * k.mesh.np};
    mgchem::InferArgs ia{};
    ia.n = k.mesh.nelements * k.mesh.np;
    ia.np = k.mesh.np;
    ia.dim = 3;
    ia.dt = dt;
    ia.q = k.q;
    ia.temp = k.temperature;
    ia.out = k.rhs;
    ia.add = 1;
    ia.prm = mlp_params;
    ia.wbf = reinterpret_cast<const __bf16*>(mlp_wbf16);
    ia.fallback = fallback_count;
    if (surrogate) ia.nm = to_norm(norm);
    st_[c] = ia;
  }
  // segments: by decreasing order, case by case
  std::memset(plan, 0, sizeof(*plan));
  const bool fused = !tuning || tuning->fused;
  const int cus = device_cus();
  for (int phase = 0; phase < 2; ++phase) {
    Seg* sg = reinterpret_cast<Seg*>(host.data() + (phase ? L.div_items : L.grad_items));
    int32_t* sb = phase ? plan->div_seg_begin : plan->grad_seg_begin;
    int32_t* sn = phase ? plan->div_seg_count : plan->grad_seg_count;
    int64_t* ib = phase ? plan->div_begin : plan->grad_begin;
    int64_t* in = phase ? plan->div_count : plan->grad_count;
    int nseg = 0;
    int64_t pos = 0;
    for (int p = 4; p >= 1; --p) {
      sb[p] = nseg;
      ib[p] = pos;
      for (int c = 0; c < ncases; ++c) {
        if (cases[c].mesh.order != p) continue;
        sg[nseg++] = Seg{c, p, pos};
        pos += case_items(cases[c].mesh, phase, surrogate);
      }
      sn[p] = nseg - sb[p];
      in[p] = pos - ib[p];
    }
    sb[0] = 0;
    sn[0] = nseg;
    ib[0] = 0;
    in[0] = pos;
  }
  for (int phase = 2; phase < 4; ++phase) {
    Seg* sg = reinterpret_cast<Seg*>(host.data() + (phase == 2 ? L.pt_items : L.src_items));
    int64_t pos = 0;
    for (int c = 0; c < ncases; ++c) {
      sg[c] = Seg{c, 0, pos};
      pos += case_items(cases[c].mesh, phase, surrogate);
    }
    (phase == 2 ? plan->pt_count : plan->src_count) = pos;
  }
  plan->dev = dev;
  plan->ncases = ncases;
  plan->fused = fused ? 1 : 0;
  plan->persistent = (tuning && tuning->persistent) ? 1 : 0;
  plan->niter = niter;
  plan->cus = cus;
  plan->surrogate = surrogate ? 1 : 0;
  plan->off_grad_tab = L.grad_tab;
  plan->off_div_tab = L.div_tab;
  plan->off_pt_tab = L.pt_tab;
  plan->off_src_tab = L.src_tab;
  plan->off_grad_items = L.grad_items;
  plan->off_div_items = L.div_items;
  plan->off_pt_items = L.pt_items;
  plan->off_src_items = L.src_items;
  for (int p = 0; p <= 4; ++p) {
    const int wg = tuning ? tuning->wg_per_cu[p] : (p == 0 ? 4 : 4);
    plan->grid_grad[p] = grid_for(plan->grad_count[p], wg, cus);
    plan->grid_div[p] = grid_for(plan->div_count[p], wg, cus);
  }
  plan->grid_pt = grid_for(plan->pt_count, 8, cus);
  // surrogate: one workgroup per 128-row item (a persistent loop over the
  // items at two workgroups per CU ran 15 % slower than the per-case
  // launches); exact rates: persistent, 8 per CU
  plan->grid_src = surrogate ? int(std::min<int64_t>(std::max<int64_t>(plan->src_count, 1),
                                                     0x7fffffff))
                             : grid_for(plan->src_count, 8, cus);
  hipStream_t stm = static_cast<hipStream_t>(stream);
  hipError_t e = hipMemcpyAsync(dev, host.data(), size_t(L.bytes), hipMemcpyHostToDevice, stm);
  if (e == hipSuccess) e = hipStreamSynchronize(stm);
  if (e != hipSuccess) return fail(MG_ERR_HIP, std::string("plan upload: ") + hipGetErrorString(e));
  return MG_OK;
}

mg_status mg_ensemble_ns_rhs(const mg_ens_plan* plan, void* stream) {
  if (!plan || !plan->dev || plan->ncases < 1) return fail(MG_ERR_ARG, "plan not prepared");
  mg_status s;
  if ((s = dg_attrs())) return s;
  hipStream_t st = static_cast<hipStream_t>(stream);
  char* base = static_cast<char*>(plan->dev);
  const auto* gt = reinterpret_cast<const mgens::KArgs*>(base + plan->off_grad_tab);
  const auto* dtab = reinterpret_cast<const mgens::KArgs*>(base + plan->off_div_tab);
  const auto* pt = reinterpret_cast<const mgens::PointCase*>(base + plan->off_pt_tab);
  const auto* stab = reinterpret_cast<const mgens::InferArgs*>(base + plan->off_src_tab);
  const auto* gi = reinterpret_cast<const Seg*>(base + plan->off_grad_items);
  const auto* di = reinterpret_cast<const Seg*>(base + plan->off_div_items);
  const auto* pi = reinterpret_cast<const Seg*>(base + plan->off_pt_items);
  const auto* si = reinterpret_cast<const Seg*>(base + plan->off_src_items);
  if (plan->pt_count > 0) {
    hipLaunchKernelGGL(mgens::ens_temperature_kernel, dim3(unsigned(plan->grid_pt)),
                       dim3(kBlock), 0, st, pt, pi, plan->ncases, plan->pt_count,
                       plan->niter);
    if ((s = check_launch("ens_temperature_kernel"))) return s;
  }
  for (int phase = 0; phase < 2; ++phase) {
    if (phase == 1 && plan->pt_count > 0) {
      // phase N between the gradients and the divergence
      hipLaunchKernelGGL(mgens::ens_node_flux_kernel, dim3(unsigned(plan->grid_pt)),
                         dim3(kBlock), 0, st, gt, pi, plan->ncases, plan->pt_count);
      if ((s = check_launch("ens_node_flux_kernel"))) return s;
    }
    const auto* tab = phase ? dtab : gt;
    const Seg* segs = phase ? di : gi;
    const int32_t* sb = phase ? plan->div_seg_begin : plan->grad_seg_begin;
    const int32_t* sn = phase ? plan->div_seg_count : plan->grad_seg_count;
    const int64_t* begin = phase ? plan->div_begin : plan->grad_begin;
    const int64_t* count = phase ? plan->div_count : plan->grad_count;
    const int32_t* grid = phase ? plan->grid_div : plan->grid_grad;
    if (plan->fused) {
      if ((s = launch_dg(phase, 0, plan->persistent, grid[0], lds_of(phase, 0), tab,
                         segs + sb[0], sn[0], begin[0], count[0], st)))
        return s;
    } else {
      for (int p = 4; p >= 1; --p)
        if (sn[p] > 0 &&
            (s = launch_dg(phase, p, plan->persistent, grid[p], lds_of(phase, p), tab,
                           segs + sb[p], sn[p], begin[p], count[p], st)))
          return s;
    }
  }
  if (plan->src_count > 0) {
    if (plan->surrogate)
      hipLaunchKernelGGL(mgens::ens_infer_kernel, dim3(unsigned(plan->grid_src)), dim3(256),
                         mgchem::kInferLds, st, stab, si, plan->ncases, plan->src_count);
    else
      hipLaunchKernelGGL(mgens::ens_exact_source_kernel, dim3(unsigned(plan->grid_src)),
                         dim3(kBlock), 0, st, stab, si, plan->ncases, plan->src_count);
    if ((s = check_launch("ensemble source"))) return s;
  }
  return MG_OK;
}

}  // extern "C"
