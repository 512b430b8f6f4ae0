// Chemistry of the uiuc_7sp mixture beyond the pointwise source: the
// constant-pressure reactor (the trajectories the learned surrogate is
// trained on; the reactor of test/test_chemistry.py:51-110) and the learned
// chemistry surrogate itself, a 10 -> 256 -> 256 -> 256 -> 7 GELU network in
// BF16 on the matrix cores (v_mfma_f32_16x16x32_bf16, FP32 accumulation):
// training (forward, backward, Adam) and the fused, gated inference that
// adds the species sources into a right-hand side.
//
// Network layout (all widths padded to the matrix-core tiles):
//   input  32 features: (log T, log p, Y_1..Y_7, dt) normalised, then zeros
//   hidden 256, 256, 256 with GELU (erf form)
//   output 32: normalised dY_k over the substep dt in 0..6, zeros after
// Parameters: one FP32 master array (W_l as [K_in][N_out] row-major, then
// b_l, layer by layer; MG_MLP_NPARAM values) and BF16 copies of every W_l
// in two layouts, Wt_l [N_out][K_in] (forward) and W_l [K_in][N_out]
// (backward), written by the Adam kernel.
#include <mutex>
#include <string>

#include "mgdg_host.h"
#include "mgchem_kernels.h"


namespace {

mg_status launch_gemm(int epi, const mgchem::GemmArgs& g, hipStream_t st) {
  if (g.m % 32 || g.n % 16 || g.k % 32 || g.ksplit < 1 || g.k % g.ksplit ||
      (g.k / g.ksplit) % 32 || g.lda < g.k || g.ldb < g.k || g.lda % 8 || g.ldb % 8)
    return fail(MG_ERR_ARG, "GEMM shape: M % 32, N % 16, (K / ksplit) % 32, lda/ldb");
  const dim3 grid(unsigned((g.m + 63) / 64), unsigned((g.n + 63) / 64),
                  unsigned(g.ksplit));
  if (epi == mgchem::kEpiStore)
    hipLaunchKernelGGL(mgchem::gemm_nt_kernel<mgchem::kEpiStore>, grid, dim3(256), 0, st, g);
  else if (epi == mgchem::kEpiGelu)
    hipLaunchKernelGGL(mgchem::gemm_nt_kernel<mgchem::kEpiGelu>, grid, dim3(256), 0, st, g);
  else
    hipLaunchKernelGGL(mgchem::gemm_nt_kernel<mgchem::kEpiBwd>, grid, dim3(256), 0, st, g);
  return check_launch("gemm_nt_kernel");
}

// Workspace of a training batch B (all offsets 256-byte aligned):
//   act1..3  BF16 [B][256] and act1..3^T, actx^T BF16 [32][B]
//   z1..3    FP32 [B][256];  out FP32 [B][32]
//   dz1..3   BF16 [B][256] and ^T;  dout BF16 [B][32] and ^T
//   gpart    FP32 [ksplit][NPARAM];  loss_part FP32 [B*16/256]
struct Work {
  __bf16 *act[3], *act_t[3], *x_t;
  float* z[3];
  float* out;
  __bf16 *dz[3], *dz_t[3], *dout, *dout_t;
  float* gpart;
  float* loss_part;
  int64_t bytes;
};

constexpr int kKsplit = 8;

Work carve(char* base, int64_t b) {
  Work w{};
  int64_t off = 0;
  auto take = [&](int64_t nbytes) {
    char* p = base ? base + off : nullptr;
    off += (nbytes + 255) / 256 * 256;
    return p;
  };
  for (int l = 0; l < 3; ++l) {
    w.act[l] = reinterpret_cast<__bf16*>(take(b * mgchem::kH * 2));
    w.act_t[l] = reinterpret_cast<__bf16*>(take(b * mgchem::kH * 2));
    w.z[l] = reinterpret_cast<float*>(take(b * mgchem::kH * 4));
    w.dz[l] = reinterpret_cast<__bf16*>(take(b * mgchem::kH * 2));
    w.dz_t[l] = reinterpret_cast<__bf16*>(take(b * mgchem::kH * 2));
  }
  w.x_t = reinterpret_cast<__bf16*>(take(b * mgchem::kIn * 2));
  w.out = reinterpret_cast<float*>(take(b * mgchem::kOut * 4));
  w.dout = reinterpret_cast<__bf16*>(take(b * mgchem::kOut * 2));
  w.dout_t = reinterpret_cast<__bf16*>(take(b * mgchem::kOut * 2));
  w.gpart = reinterpret_cast<float*>(take(int64_t(kKsplit) * mgchem::kNParam * 4));
  w.loss_part = reinterpret_cast<float*>(take((b * mgchem::kOut + 255) / 256 * 4));
  w.bytes = off;
  return w;
}

__global__ __launch_bounds__(256) void transpose_x_kernel(int64_t b, const __bf16* x,
                                                          __bf16* xt) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= b * mgchem::kIn) return;
  const int64_t row = i / mgchem::kIn;
  const int64_t col = i - row * mgchem::kIn;
  xt[col * b + row] = x[i];
}

mg_status check_batch(int64_t b) {
  if (b < 64 || b % 64) return fail(MG_ERR_ARG, "batch must be a positive multiple of 64");
  if (b % (32 * kKsplit)) return fail(MG_ERR_ARG, "batch must be a multiple of 256");
  if (b > (int64_t(1) << 24)) return fail(MG_ERR_ARG, "batch above 2^24");
  return MG_OK;
}

mg_status forward(int64_t b, const __bf16* x, const float* prm, const __bf16* wbf,
                  const Work& w, hipStream_t st) {
  using namespace mgchem;
  mg_status s;
  const __bf16* in = x;
  int64_t kin = kIn;
  const int64_t boff[3] = {kOffB1, kOffB2, kOffB3};
  const int64_t wof[3] = {kBfW1, kBfW2, kBfW3};
  for (int l = 0; l < 3; ++l) {
    GemmArgs g{};
    g.a = in;
    g.lda = kin;
    g.bt = wbf + wof[l];
    g.ldb = kin;
    g.m = b;
    g.n = kH;
    g.k = kin;
    g.ksplit = 1;
    g.bias = prm + boff[l];
    g.zout = w.z[l];
    g.out = w.act[l];
    g.out_t = w.act_t[l];
    if ((s = launch_gemm(kEpiGelu, g, st))) return s;
    in = w.act[l];
    kin = kH;
  }
  GemmArgs g{};
  g.a = in;
  g.lda = kH;
  g.bt = wbf + kBfW4;
  g.ldb = kH;
  g.m = b;
  g.n = kOut;
  g.k = kH;
  g.ksplit = 1;
  g.bias = prm + kOffB4;
  g.c = w.out;
  g.ldc = kOut;
  g.cslice = 0;
  return launch_gemm(kEpiStore, g, st);
}

}  // namespace

extern "C" {

int64_t mg_mlp_param_count(void) { return mgchem::kNParam; }

int64_t mg_mlp_bf16_count(void) { return 2 * mgchem::kBfHalf; }

int64_t mg_mlp_workspace_bytes(int64_t batch) {
  if (check_batch(batch)) return -1;
  return carve(nullptr, batch).bytes;
}

mg_status mg_reactor_trajectories(int64_t n, const double* t0, const double* p,
                                  const double* y0, double dt, int64_t nmax, double t_rise,
                                  int32_t nsamp, double* t_ign, double* states, double* dy,
                                  void* stream) {
  mg_status s;
  if ((s = ensure_mech())) return s;
  if (n < 0 || !t0 || !p || !y0 || !t_ign || !(dt > 0.0) || nmax < 1 || nsamp < 0)
    return fail(MG_ERR_ARG, "bad arguments");
  if (nsamp > 0 && (!states || !dy)) return fail(MG_ERR_ARG, "states/dy NULL");
  if (n == 0) return MG_OK;
  hipLaunchKernelGGL(mgchem::reactor_kernel, dim3(unsigned((n + 63) / 64)), dim3(64), 0,
                     static_cast<hipStream_t>(stream), n, t0, p, y0, dt, nmax, t_rise,
                     int(nsamp), t_ign, states, dy);
  return check_launch("reactor_kernel");
}

mg_status mg_chem_substep(int64_t n, const double* states, double dt, int32_t nsub,
                          double* dy, void* stream) {
  mg_status s;
  if ((s = ensure_mech())) return s;
  if (n < 0 || !states || !dy || !(dt > 0.0) || nsub < 1)
    return fail(MG_ERR_ARG, "bad arguments");
  if (n == 0) return MG_OK;
  hipLaunchKernelGGL(mgchem::substep_kernel, dim3(unsigned((n + 63) / 64)), dim3(64), 0,
                     static_cast<hipStream_t>(stream), n, states, dt, int(nsub), dy);
  return check_launch("substep_kernel");
}

mg_status mg_chem_features(int64_t n, const double* states, const double* dy,
                           const mg_chem_norm* norm, uint16_t* x, float* target,
                           void* stream) {
  mg_status s;
  if ((s = check_norm(norm))) return s;
  if (n < 0 || !states || !x || (dy && !target)) return fail(MG_ERR_ARG, "bad arguments");
  if (n == 0) return MG_OK;
  hipLaunchKernelGGL(mgchem::prepare_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), n, states, dy, to_norm(norm),
                     reinterpret_cast<__bf16*>(x), target);
  return check_launch("prepare_kernel");
}

mg_status mg_mlp_forward(int64_t batch, const uint16_t* x, const float* params,
                         const uint16_t* wbf16, void* work, float* out, void* stream) {
  mg_status s;
  if ((s = check_batch(batch))) return s;
  if (!x || !params || !wbf16 || !work) return fail(MG_ERR_ARG, "NULL argument");
  Work w = carve(static_cast<char*>(work), batch);
  hipStream_t st = static_cast<hipStream_t>(stream);
  if ((s = forward(batch, reinterpret_cast<const __bf16*>(x), params,
                   reinterpret_cast<const __bf16*>(wbf16), w, st)))
    return s;
  if (out && out != w.out) {
    hipError_t e = hipMemcpyAsync(out, w.out, sizeof(float) * batch * mgchem::kOut,
                                  hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess) return fail(MG_ERR_HIP, hipGetErrorString(e));
  }
  return MG_OK;
}

mg_status mg_mlp_backward(int64_t batch, const uint16_t* x, const float* target,
                          const float* params, const uint16_t* wbf16, void* work,
                          float* loss_part, void* stream) {
  using namespace mgchem;
  mg_status s;
  if ((s = check_batch(batch))) return s;
  if (!x || !target || !params || !wbf16 || !work) return fail(MG_ERR_ARG, "NULL argument");
  Work w = carve(static_cast<char*>(work), batch);
  hipStream_t st = static_cast<hipStream_t>(stream);
  const __bf16* wb = reinterpret_cast<const __bf16*>(wbf16) + kBfHalf;  // W [K][N]
  const int64_t nl = (batch * kOut + 255) / 256;
  hipLaunchKernelGGL(loss_kernel, dim3(unsigned(nl)), dim3(256), 0, st, batch, w.out, target,
                     w.dout, w.dout_t, loss_part ? loss_part : w.loss_part);
  if ((s = check_launch("loss_kernel"))) return s;
  hipLaunchKernelGGL(transpose_x_kernel, dim3(unsigned((batch * kIn + 255) / 256)),
                     dim3(256), 0, st, batch, reinterpret_cast<const __bf16*>(x), w.x_t);
  if ((s = check_launch("transpose_x_kernel"))) return s;
  // layer l (0..3): input act_{l-1} (x for l = 0), output grad dz_l
  const __bf16* in_t[4] = {w.x_t, w.act_t[0], w.act_t[1], w.act_t[2]};
  const int kin[4] = {kIn, kH, kH, kH};
  const int nout[4] = {kH, kH, kH, kOut};
  const int64_t woff[4] = {kOffW1, kOffW2, kOffW3, kOffW4};
  const int64_t boff[4] = {kOffB1, kOffB2, kOffB3, kOffB4};
  const int64_t wbo[4] = {kBfW1, kBfW2, kBfW3, kBfW4};
  for (int l = 3; l >= 0; --l) {
    const __bf16* dzl = l == 3 ? w.dout : w.dz[l];
    const __bf16* dzl_t = l == 3 ? w.dout_t : w.dz_t[l];
    // dW_l [K_in][N_out] = act_{l-1}^T dz_l, the batch split in kKsplit
    // slices whose partial sums land kNParam apart (summed by Adam)
    GemmArgs g{};
    g.a = in_t[l];
    g.lda = batch;
    g.bt = dzl_t;
    g.ldb = batch;
    g.m = kin[l];
    g.n = nout[l];
    g.k = batch;
    g.ksplit = kKsplit;
    g.c = w.gpart + woff[l];
    g.ldc = nout[l];
    g.cslice = kNParam;
    if ((s = launch_gemm(kEpiStore, g, st))) return s;
    hipLaunchKernelGGL(bias_grad_kernel, dim3(unsigned(nout[l])), dim3(256), 0, st, batch,
                       dzl_t, w.gpart, boff[l], kKsplit);
    if ((s = check_launch("bias_grad_kernel"))) return s;
    if (l == 0) break;
    // dz_{l-1} = (dz_l W_l^T) * gelu'(z_{l-1})
    GemmArgs gb{};
    gb.a = dzl;
    gb.lda = nout[l];
    gb.bt = wb + wbo[l];
    gb.ldb = nout[l];
    gb.m = batch;
    gb.n = kin[l];
    gb.k = nout[l];
    gb.ksplit = 1;
    gb.zpre = w.z[l - 1];
    gb.out = w.dz[l - 1];
    gb.out_t = w.dz_t[l - 1];
    if ((s = launch_gemm(kEpiBwd, gb, st))) return s;
  }
  return MG_OK;
}

mg_status mg_mlp_adam(float* params, float* m, float* v, const void* work, int64_t batch,
                      int32_t step, float lr, float beta1, float beta2, float eps,
                      uint16_t* wbf16, float* grad_out, void* stream) {
  mg_status s;
  if ((s = check_batch(batch))) return s;
  if (!params || !m || !v || !work || !wbf16) return fail(MG_ERR_ARG, "NULL argument");
  if (step < 1) return fail(MG_ERR_ARG, "step counts from 1");
  Work w = carve(static_cast<char*>(const_cast<void*>(work)), batch);
  const float bc1 = 1.0f - powf(beta1, float(step));
  const float bc2 = 1.0f - powf(beta2, float(step));
  hipLaunchKernelGGL(mgchem::adam_kernel, dim3(unsigned((mgchem::kNParam + 255) / 256)),
                     dim3(256), 0, static_cast<hipStream_t>(stream), params, m, v, w.gpart,
                     kKsplit, lr, beta1, beta2, eps, bc1, bc2,
                     reinterpret_cast<__bf16*>(wbf16), grad_out);
  return check_launch("adam_kernel");
}

mg_status mg_mlp_sync_bf16(const float* params, uint16_t* wbf16, void* stream) {
  if (!params || !wbf16) return fail(MG_ERR_ARG, "NULL argument");
  // an Adam pass with lr = 0 and no gradient slices only re-emits the copies
  hipLaunchKernelGGL(mgchem::adam_kernel, dim3(unsigned((mgchem::kNParam + 255) / 256)),
                     dim3(256), 0, static_cast<hipStream_t>(stream),
                     const_cast<float*>(params), nullptr, nullptr, nullptr, 0, 0.0f, 0.0f,
                     0.0f, 0.0f, 1.0f, 1.0f, reinterpret_cast<__bf16*>(wbf16), nullptr);
  return check_launch("adam_kernel");
}

mg_status mg_chem_surrogate_source(const mg_gas* gas, int32_t dim, int32_t np, int64_t n,
                                   const double* q, const double* temperature, double dt,
                                   const float* params, const uint16_t* wbf16,
                                   const mg_chem_norm* norm, int32_t add, double* out,
                                   int32_t* fallback_count, void* stream) {
  mg_status s;
  if ((s = check_gas(gas, false))) return s;
  if (gas->eos_kind != 1) return fail(MG_ERR_ARG, "mixture EOS required");
  if ((s = ensure_mech())) return s;
  if ((s = check_norm(norm))) return s;
  if (dim != 2 && dim != 3) return fail(MG_ERR_ARG, "dim");
  if (np < 1 || np > mgchem::kRows) return fail(MG_ERR_ARG, "np must be in 1..128");
  if (n < 0 || n % np) return fail(MG_ERR_ARG, "n must be a multiple of np");
  if (!q || !temperature || !params || !wbf16 || !out || !(dt > 0.0))
    return fail(MG_ERR_ARG, "bad arguments");
  if (n == 0) return MG_OK;
  mgchem::InferArgs a{};
  a.n = n;
  a.np = np;
  a.dim = dim;
  a.dt = dt;
  a.q = q;
  a.temp = temperature;
  a.out = out;
  a.add = add ? 1 : 0;
  a.prm = params;
  a.wbf = reinterpret_cast<const __bf16*>(wbf16);
  a.fallback = fallback_count;
  a.nm = to_norm(norm);
  const int rows = (mgchem::kRows / np) * np;
  const int64_t nblk = (n + rows - 1) / rows;
  constexpr size_t lds = mgchem::kInferLds;
  static bool attr = false;
  if (!attr) {
    if ((s = set_lds(mgchem::infer_kernel, lds))) return s;
    attr = true;
  }
  hipLaunchKernelGGL(mgchem::infer_kernel, dim3(unsigned(nblk)), dim3(256), lds,
                     static_cast<hipStream_t>(stream), a);
  return check_launch("infer_kernel");
}

}  // extern "C"
