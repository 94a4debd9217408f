// C-ABI of the convolutional closure network (BF16 MFMA convolutions,
// gradients, Adam) and of the learned-closure LES term.
#include "fsx_internal.h"

#include <mutex>

namespace {

bool cnn_channels_ok(int c) { return c == 8 || c == 32; }

int cnn_grid(long long n) {
  long long b = (n + 255) / 256;
  return (int)std::max<long long>(1, std::min<long long>(b, 4096));
}

// fuse: 0 -> FP32 pre-activations y; 1 -> BF16 GELU activations a (32
// channels, co_out = 32); 2 -> both; 3 -> a = BF16((conv + bias) GELU'(y))
// with y the layer's pre-activations (read; co_out = 32)
template <int CI, int FUSE, bool XF32>
void conv_fwd_launch(unsigned nblk, size_t lds, hipStream_t s, const uint16_t* x,
                     const uint16_t* w, const float* bias, float* y, int H, int W, int co_out,
                     int R, uint16_t* a) {
  // 4-row blocks, more than two per CU: the persistent form (weights staged
  // once per workgroup, patch rows in a ring), one workgroup per CU;
  // blocks of 2 rows: 8 waves (more tiles in flight per block; the LDS
  // allows one block per CU), one-row blocks: 4 waves
  static int ncu = 0;
  if (!ncu) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        ncu <= 0)
      ncu = 256;
  }
  if (R == kCnnRows && nblk > 2u * (unsigned)ncu) {
    const int B = (int)(nblk / (unsigned)(H / R));
    const unsigned grid = (unsigned)(ncu >= 8 ? ncu / 8 * 8 : ncu);
    hipLaunchKernelGGL((k_conv_fwd_run<CI, FUSE, XF32>), dim3(grid), dim3(512), lds, s, x, w,
                       bias, y, B, H, W, co_out, a);
  } else if (R >= 2)
    hipLaunchKernelGGL((k_conv_fwd<CI, FUSE, XF32, 512>), dim3(nblk), dim3(512), lds, s, x, w,
                       bias, y, H, W, co_out, R, a);
  else
    hipLaunchKernelGGL((k_conv_fwd<CI, FUSE, XF32, 256>), dim3(nblk), dim3(256), lds, s, x, w,
                       bias, y, H, W, co_out, R, a);
}

// x_f32: x is one FP32 channel (ci = 8 layer: padded while staged)
int cnn_conv_fwd(const void* x, int B, int H, int W, int ci, const void* w, const float* bias,
                 int co_out, float* y, hipStream_t s, int fuse = 0, void* a = nullptr,
                 bool x_f32 = false) {
  if (!x || !w || !bias || B < 1 || fuse < 0 || fuse > 3) return fail(FSX_ERR_ARG, "bad argument");
  if ((fuse != 1 && !y) || (fuse != 0 && !a)) return fail(FSX_ERR_ARG, "bad argument");
  if (!cnn_channels_ok(ci)) return fail(FSX_ERR_ARG, "input channels must be 8 or 32");
  if (co_out < 1 || co_out > 32) return fail(FSX_ERR_ARG, "output channels must be 1..32");
  if (fuse && co_out != 32) return fail(FSX_ERR_ARG, "fused activations need 32 output channels");
  if (x_f32 && (ci != 8 || fuse != 1))
    return fail(FSX_ERR_ARG, "an FP32 input is one channel (padded to 8) of a hidden layer");
  if (H < kCnnRows || H % kCnnRows || W < 32 || W % 32)
    return fail(FSX_ERR_ARG, "H must be a multiple of 4 and W a multiple of 32");
  // rows per workgroup: 4 when there are enough workgroups, fewer for
  // small batches (more workgroups, smaller patches)
  int R = kCnnRows;
  while (R > 1 && (long long)B * (H / R) < 512) R >>= 1;
  const size_t lds = conv_fwd_lds(ci, W, R);
  if (lds > kLdsMax) return fail(FSX_ERR_ARG, "image too wide for the convolution tile");
  const unsigned nblk = (unsigned)(B * (H / R));
  const auto* xi = reinterpret_cast<const uint16_t*>(x);
  const auto* wi = reinterpret_cast<const uint16_t*>(w);
  auto* ai = reinterpret_cast<uint16_t*>(a);
  int rc;
#define FSX_CONV_FWD(CI_, F_, XF_)                                                      \
  if (ci == CI_ && fuse == F_ && x_f32 == XF_) {                                          \
    rc = R >= 2 ? set_lds(k_conv_fwd<CI_, F_, XF_, 512>, lds)                             \
                : set_lds(k_conv_fwd<CI_, F_, XF_, 256>, lds);                            \
    if (!rc && R == kCnnRows) rc = set_lds(k_conv_fwd_run<CI_, F_, XF_>, lds);            \
    if (rc) return rc;                                                                    \
    if ((rc = tm_pre(s))) return rc;                                                      \
    conv_fwd_launch<CI_, F_, XF_>(nblk, lds, s, xi, wi, bias, y, H, W, co_out, R, ai);    \
  }
  FSX_CONV_FWD(8, 0, false)
  FSX_CONV_FWD(8, 1, false)
  FSX_CONV_FWD(8, 2, false)
  FSX_CONV_FWD(8, 1, true)
  FSX_CONV_FWD(32, 0, false)
  FSX_CONV_FWD(32, 1, false)
  FSX_CONV_FWD(32, 2, false)
  FSX_CONV_FWD(8, 3, false)
  FSX_CONV_FWD(32, 3, false)
#undef FSX_CONV_FWD
  FSX_HIP(hipGetLastError());
  return tm_post(s);
}

int cnn_act(const float* y, long long npix, int cin, int cpad, int use_gelu, void* a,
            hipStream_t s) {
  if (!y || !a || npix < 1 || cin < 1 || cpad < cin) return fail(FSX_ERR_ARG, "bad argument");
  hipLaunchKernelGGL(k_cnn_act, dim3(cnn_grid(npix * cpad)), dim3(256), 0, s, y, npix, cin,
                     cpad, use_gelu, reinterpret_cast<uint16_t*>(a));
  FSX_HIP(hipGetLastError());
  return 0;
}

// two BF16 activation buffers of 32 channels
long long cnn_forward_work(int B, int H, int W) { return 2LL * B * H * W * 32 * 2; }

// Network forward: hidden layers write their BF16 GELU activations from the
// convolution epilogue (ping-pong between the two buffers of work), the last
// layer writes its FP32 outputs to y.
// x_f32: x is one FP32 channel (the first layer pads it to 8 while staging)
int cnn_forward(int nl, const int* cin, const int* cout, const void* const* w,
                const float* const* b, const void* x, int B, int H, int W, float* y,
                void* work, hipStream_t s, bool x_f32 = false) {
  const long long npix = (long long)B * H * W;
  uint16_t* act[2] = {reinterpret_cast<uint16_t*>(work),
                      reinterpret_cast<uint16_t*>(work) + npix * 32};
  const void* in = x;
  for (int l = 0; l < nl; ++l) {
    const bool last = l == nl - 1;
    if (!cnn_channels_ok(cin[l])) return fail(FSX_ERR_ARG, "layer input channels must be 8 or 32");
    if (!w[l] || !b[l]) return fail(FSX_ERR_ARG, "bad argument");
    if (!last && cout[l] != 32) return fail(FSX_ERR_ARG, "hidden layers have 32 channels");
    // timing: hidden 32 -> 32 layers as their own kind, work in flops
    if (l > 0 && !last)
      note(K_CNN, 2.0 * npix * 32 * kCnnTaps * 32);
    else
      note(K_OTHER, 0.0);
    if (x_f32 && l == 0 && last) return fail(FSX_ERR_ARG, "an FP32 input needs a hidden first layer");
    const int rc = last ? cnn_conv_fwd(in, B, H, W, cin[l], w[l], b[l], cout[l], y, s)
                        : cnn_conv_fwd(in, B, H, W, cin[l], w[l], b[l], 32, nullptr, s, 1,
                                       act[l & 1], x_f32 && l == 0);
    if (rc) return rc;
    in = act[l & 1];
  }
  return 0;
}

// Weight and bias gradients into dw / dbias through the partials buffer
// part (wgrad_groups(B, H) x (32 x 25 ci + 32) floats); arguments checked
// by the caller.
int cnn_wgrad(const void* x, int ci, const void* dy, int co_in, int B, int H, int W, float* dw,
              float* dbias, float* part, hipStream_t s) {
  // row blocks of kWgRowsMax rows where their LDS fits and H divides
  const bool wide = conv_wgrad_lds(ci, W, kWgRowsMax) <= kLdsMax && H % kWgRowsMax == 0;
  const int rows = wide ? kWgRowsMax : kWgRows;
  const size_t lds = conv_wgrad_lds(ci, W, rows);
  if (lds > kLdsMax) return fail(FSX_ERR_ARG, "image too wide for the weight-gradient tile");
  const int groups = wgrad_groups(B, H, rows);
  const auto* xi = reinterpret_cast<const uint16_t*>(x);
  const auto* di = reinterpret_cast<const uint16_t*>(dy);
  int rc;
#define FSX_WG(CI_, CO_)                                                                    \
  if (wide) {                                                                               \
    if ((rc = set_lds(k_conv_wgrad<CI_, CO_, kWgRowsMax>, lds))) return rc;                 \
    hipLaunchKernelGGL((k_conv_wgrad<CI_, CO_, kWgRowsMax>), dim3(groups), dim3(kWgThreads), \
                       lds, s, xi, di, part, B, H, W);                                      \
  } else {                                                                                  \
    if ((rc = set_lds(k_conv_wgrad<CI_, CO_, kWgRows>, lds))) return rc;                    \
    hipLaunchKernelGGL((k_conv_wgrad<CI_, CO_, kWgRows>), dim3(groups), dim3(kWgThreads),   \
                       lds, s, xi, di, part, B, H, W);                                      \
  }
  if (ci == 8 && co_in == 8) {
    FSX_WG(8, 8)
  } else if (ci == 8) {
    FSX_WG(8, 32)
  } else if (co_in == 8) {
    FSX_WG(32, 8)
  } else {
    FSX_WG(32, 32)
  }
#undef FSX_WG
  FSX_HIP(hipGetLastError());
  const int n = 32 * kCnnTaps * ci;
  const int nb1 = (n + 63) / 64;  // then one workgroup for the 32 bias sums
  hipLaunchKernelGGL(k_cnn_reduce, dim3(nb1 + 1), dim3(256), 0, s, part, groups, n, dw, nb1,
                     part + (size_t)groups * n, 32, dbias);
  FSX_HIP(hipGetLastError());
  return 0;
}

// Scratch of fsx_cnn_conv_wgrad_noalloc: the partials buffer of cnn_wgrad at
// its largest (kWgGroups workgroups, 32 input channels), a device array in
// the code object, so that the entry point allocates nothing.  Each device
// holds its own copy; host code reaches it only through the address
// hipGetSymbolAddress returns on that device (the array's name is a host
// stand-in there).
constexpr long long kWgradScratchFloats = (long long)kWgGroups * (32LL * kCnnTaps * 32 + 32);
static __device__ __align__(256) float scratch[kWgradScratchFloats];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float* g_scratch_ptr[kMaxDevices] = {};

// Address of the current device's scratch, resolved once per device.
int wgrad_scratch(float** out) {
  int dev = 0;
  FSX_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(FSX_ERR_ARG, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    FSX_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return fail(FSX_ERR_HIP, "no address for the weight-gradient scratch");
    g_scratch_ptr[dev] = reinterpret_cast<float*>(ptr);
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

// workspace of fsx_ns2d_closure_term: vorticity and network output (FP32),
// the network's activations
long long closure_work(const fsx_plan_s* p) {
  const long long npix = (long long)p->ny * p->nx;
  return npix * (4 + 4) + cnn_forward_work(1, p->ny, p->nx);
}

}  // namespace

extern "C" {

int fsx_cnn_probe(const void* A, const void* B, float* D, void* stream) {
  if (!A || !B || !D) return fail(FSX_ERR_ARG, "bad argument");
  hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, (hipStream_t)stream,
                     reinterpret_cast<const uint16_t*>(A), reinterpret_cast<const uint16_t*>(B),
                     D);
  FSX_HIP(hipGetLastError());
  return 0;
}

int fsx_cnn_probe16(const void* A, const void* B, float* D, void* stream) {
  if (!A || !B || !D) return fail(FSX_ERR_ARG, "bad argument");
  hipLaunchKernelGGL(k_mfma16_probe, dim3(1), dim3(64), 0, (hipStream_t)stream,
                     reinterpret_cast<const uint16_t*>(A), reinterpret_cast<const uint16_t*>(B),
                     D);
  FSX_HIP(hipGetLastError());
  return 0;
}

int fsx_cnn_conv_fwd(const void* x, int B, int H, int W, int ci, const void* w,
                     const float* bias, int co_out, float* y, void* stream) {
  return cnn_conv_fwd(x, B, H, W, ci, w, bias, co_out, y, (hipStream_t)stream);
}

int fsx_cnn_conv_fwd_act(const void* x, int B, int H, int W, int ci, const void* w,
                         const float* bias, float* y, void* a, void* stream) {
  return cnn_conv_fwd(x, B, H, W, ci, w, bias, 32, y, (hipStream_t)stream, y ? 2 : 1, a);
}

int fsx_cnn_conv_bwd_act(const void* x, int B, int H, int W, int ci, const void* w,
                         const float* bias, const float* pre, void* d, void* stream) {
  if (!pre || !d) return fail(FSX_ERR_ARG, "bad argument");
  return cnn_conv_fwd(x, B, H, W, ci, w, bias, 32, const_cast<float*>(pre),
                      (hipStream_t)stream, 3, d);
}

int fsx_cnn_conv_flip(const void* w, int ci, int co, int co_pad, void* wt, void* stream) {
  if (!w || !wt || ci < 1 || ci > 32 || co < 1 || co > co_pad || !cnn_channels_ok(co_pad))
    return fail(FSX_ERR_ARG, "bad argument");
  hipLaunchKernelGGL(k_conv_flip, dim3(cnn_grid(32LL * kCnnTaps * co_pad)), dim3(256), 0,
                     (hipStream_t)stream, reinterpret_cast<const uint16_t*>(w), ci, co, co_pad,
                     reinterpret_cast<uint16_t*>(wt));
  FSX_HIP(hipGetLastError());
  return 0;
}

int fsx_cnn_wgrad_workspace(int B, int H, int ci, long long* bytes) {
  if (!bytes || B < 1 || H < kCnnRows || !cnn_channels_ok(ci)) return fail(FSX_ERR_ARG, "bad argument");
  const long long groups = wgrad_groups(B, H);
  // weight partials, then bias partials
  *bytes = groups * 32LL * kCnnTaps * ci * 4 + groups * 32LL * 4;
  return 0;
}

int fsx_cnn_conv_wgrad(const void* x, int ci, const void* dy, int co_in, int B, int H, int W,
                       float* dw, float* dbias, void* work, long long work_bytes,
                       void* stream) {
  if (!x || !dy || !dw || !dbias || !work || B < 1) return fail(FSX_ERR_ARG, "bad argument");
  if (!cnn_channels_ok(ci) || !cnn_channels_ok(co_in))
    return fail(FSX_ERR_ARG, "channels must be 8 or 32");
  if (H < kCnnRows || H % kCnnRows || W < 8 || W % 8)
    return fail(FSX_ERR_ARG, "H must be a multiple of 4 and W of 8");
  long long need = 0;
  fsx_cnn_wgrad_workspace(B, H, ci, &need);
  if (work_bytes < need) return fail(FSX_ERR_ARG, "workspace too small");
  return cnn_wgrad(x, ci, dy, co_in, B, H, W, dw, dbias, reinterpret_cast<float*>(work),
                   (hipStream_t)stream);
}

int fsx_cnn_conv_wgrad_noalloc(const void* x, int ci, const void* dy, int co_in, int B, int H,
                               int W, float* dw, float* dbias, void* stream) {
  if (!x || !dy || !dw || !dbias || B < 1) return fail(FSX_ERR_ARG, "bad argument");
  if (!cnn_channels_ok(ci) || !cnn_channels_ok(co_in))
    return fail(FSX_ERR_ARG, "channels must be 8 or 32");
  if (H < kCnnRows || H % kCnnRows || W < 8 || W % 8)
    return fail(FSX_ERR_ARG, "H must be a multiple of 4 and W of 8");
  long long need = 0;
  fsx_cnn_wgrad_workspace(B, H, ci, &need);
  if (need > kWgradScratchFloats * (long long)sizeof(float))
    return fail(FSX_ERR_ARG, "partials exceed the weight-gradient scratch");
  float* part = nullptr;
  int rc = wgrad_scratch(&part);
  if (rc) return rc;
  return cnn_wgrad(x, ci, dy, co_in, B, H, W, dw, dbias, part, (hipStream_t)stream);
}

int fsx_cnn_act(const float* y, long long npix, int cin, int cpad, int use_gelu, void* a,
                void* stream) {
  return cnn_act(y, npix, cin, cpad, use_gelu, a, (hipStream_t)stream);
}

int fsx_cnn_act_bwd(const float* g, int gs, const float* y, long long npix, int c, void* d,
                    void* stream) {
  if (!g || !y || !d || npix < 1 || c < 1 || gs < c) return fail(FSX_ERR_ARG, "bad argument");
  hipLaunchKernelGGL(k_cnn_act_bwd, dim3(cnn_grid(npix * c)), dim3(256), 0,
                     (hipStream_t)stream, g, gs, y, npix, c, reinterpret_cast<uint16_t*>(d));
  FSX_HIP(hipGetLastError());
  return 0;
}

int fsx_cnn_mse(const float* out, const float* target, long long npix, int cpad, void* d,
                double* work, double* host_loss, void* stream) {
  if (!out || !target || !d || !work || npix < 1 || cpad < 1) return fail(FSX_ERR_ARG, "bad argument");
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_cnn_mse, dim3(256), dim3(256), 0, s, out, target, npix, cpad,
                     reinterpret_cast<uint16_t*>(d), work);
  FSX_HIP(hipGetLastError());
  if (host_loss) {
    double h[256];
    FSX_HIP(hipMemcpyAsync(h, work, sizeof(h), hipMemcpyDeviceToHost, s));
    FSX_HIP(hipStreamSynchronize(s));
    double t = 0.0;
    for (int i = 0; i < 256; ++i) t += h[i];
    *host_loss = t;
  }
  return 0;
}

int fsx_cnn_adam(float* w, const float* g, float* m1, float* m2, int n, double lr, double b1,
                 double b2, double eps, int step, void* wb, void* stream) {
  if (!w || !g || !m1 || !m2 || n < 1 || step < 1) return fail(FSX_ERR_ARG, "bad argument");
  const float c1 = (float)(1.0 - std::pow(b1, step)), c2 = (float)(1.0 - std::pow(b2, step));
  hipLaunchKernelGGL(k_cnn_adam, dim3(cnn_grid(n)), dim3(256), 0, (hipStream_t)stream, w, g,
                     m1, m2, n, (float)lr, (float)b1, (float)b2, (float)eps, c1, c2,
                     reinterpret_cast<uint16_t*>(wb));
  FSX_HIP(hipGetLastError());
  return 0;
}

int fsx_cnn_adam_multi(int count, float* const* w, const float* const* g, float* const* m1,
                       float* const* m2, const int* n, void* const* wb, double lr, double b1,
                       double b2, double eps, int step, void* stream) {
  if (count < 1 || count > kAdamTensors || !w || !g || !m1 || !m2 || !n || step < 1)
    return fail(FSX_ERR_ARG, "bad argument (1 to 16 tensors)");
  AdamTable tb{};
  tb.count = count;
  tb.blk[0] = 0;
  for (int t = 0; t < count; ++t) {
    if (!w[t] || !g[t] || !m1[t] || !m2[t] || n[t] < 1) return fail(FSX_ERR_ARG, "bad argument");
    tb.w[t] = w[t];
    tb.g[t] = g[t];
    tb.m1[t] = m1[t];
    tb.m2[t] = m2[t];
    tb.wb[t] = wb ? reinterpret_cast<uint16_t*>(wb[t]) : nullptr;
    tb.n[t] = n[t];
    tb.blk[t + 1] = tb.blk[t] + cnn_grid(n[t]);  // k_cnn_adam's grid
  }
  const float c1 = (float)(1.0 - std::pow(b1, step)), c2 = (float)(1.0 - std::pow(b2, step));
  hipLaunchKernelGGL(k_cnn_adam_multi, dim3(tb.blk[count]), dim3(256), 0, (hipStream_t)stream,
                     tb, (float)lr, (float)b1, (float)b2, (float)eps, c1, c2);
  FSX_HIP(hipGetLastError());
  return 0;
}

int fsx_cnn_to_bf16(const float* in, int n, void* out, void* stream) {
  if (!in || !out || n < 1) return fail(FSX_ERR_ARG, "bad argument");
  hipLaunchKernelGGL(k_cnn_to_bf16, dim3(cnn_grid(n)), dim3(256), 0, (hipStream_t)stream, in,
                     n, reinterpret_cast<uint16_t*>(out));
  FSX_HIP(hipGetLastError());
  return 0;
}

int fsx_ns2d_closure_workspace(fsx_plan p, long long* bytes) {
  if (!p || !bytes) return fail(FSX_ERR_ARG, "bad argument");
  if (p->dim != 2) return fail(FSX_ERR_ARG, "the closure term needs a 2D plan");
  *bytes = closure_work(p);
  return 0;
}

int fsx_ns2d_closure_term(fsx_plan p, int nl, const int* cin, const int* cout,
                          const void* const* w, const float* const* b, const void* rot_fft,
                          void* out_fft, void* work, long long work_bytes, void* stream) {
  PlanScope scope_(p);
  if (!p || nl < 1 || nl > 16 || !cin || !cout || !w || !b || !rot_fft || !out_fft || !work)
    return fail(FSX_ERR_ARG, "bad argument");
  if (p->dim != 2) return fail(FSX_ERR_ARG, "the closure term needs a 2D plan");
  if (p->prec != 4) return fail(FSX_ERR_ARG, "the learned closure runs on float32 plans");
  if (cin[0] != 8 || cout[nl - 1] != 1)
    return fail(FSX_ERR_ARG, "the network maps 1 input channel (padded to 8) to 1 output");
  if (work_bytes < closure_work(p)) return fail(FSX_ERR_ARG, "workspace too small");
  hipStream_t s = (hipStream_t)stream;
  const long long npix = (long long)p->ny * p->nx;
  char* wb = reinterpret_cast<char*>(work);
  float* X = reinterpret_cast<float*>(wb);
  float* Y = reinterpret_cast<float*>(wb + npix * 4);
  void* fw = wb + npix * 8;
  int rc = do_ifft<float>(p, rot_fft, X, 1, s);
  if (!rc) rc = cnn_forward(nl, cin, cout, w, b, X, 1, p->ny, p->nx, Y, fw, s, true);
  if (!rc) rc = do_fft<float>(p, Y, out_fft, 1, s);
  if (!rc) rc = fsx_dealias(p, out_fft, 1, 0, stream);
  return rc;
}

int fsx_cnn_forward(int nl, const int* cin, const int* cout, const void* const* w,
                    const float* const* b, const void* x, int B, int H, int W, float* y,
                    void* work, long long work_bytes, void* stream) {
  if (nl < 1 || nl > 16 || !cin || !cout || !w || !b || !x || !y || !work)
    return fail(FSX_ERR_ARG, "bad argument");
  hipStream_t s = (hipStream_t)stream;
  if (work_bytes < cnn_forward_work(B, H, W)) return fail(FSX_ERR_ARG, "workspace too small");
  return cnn_forward(nl, cin, cout, w, b, x, B, H, W, y, work, s);
}

}  // extern "C"
