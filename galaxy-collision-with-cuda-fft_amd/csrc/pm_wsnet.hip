// The warm-start network: a 4-level 3D convolutional encoder-decoder (U-Net) in BF16 on the
// matrix cores, with its own forward and backward passes built from the pm_conv primitives.
//
// Levels l = 0..3 run at resolution R_l = res / 2^l with C_l channels (default 16, 32, 64,
// 128). Layers, in parameter order:
//    0 e0a  in(8)        -> C0   @R0     7 e3b  C3 -> C3 @R3 (bottom)
//    1 e0b  C0 -> C0     @R0     8 d2a  up(e3b) ++ e2b -> C2 @R2
//    2 e1a  pool(e0b)    -> C1   9 d2b  C2 -> C2
//    3 e1b  C1 -> C1            10 d1a  up(d2b) ++ e1b -> C1 @R1
//    4 e2a  pool(e1b)    -> C2  11 d1b  C1 -> C1
//    5 e2b  C2 -> C2            12 d0a  up(d1b) ++ e0b -> C0 @R0
//    6 e3a  pool(e2b)    -> C3  13 d0b  C0 -> C0
//                               14 out  C0 -> 1 (no activation, float32 output)
// Every hidden layer is conv3x3x3 + bias + leaky ReLU; "pool" is 2x2x2 averaging, "up"
// nearest-neighbour upsampling (fused into the next convolution's loads), "++" channel
// concatenation (also fused). Parameters are one float32 vector: per layer the weight in
// torch layout [cout][cin][3][3][3] followed by the bias [cout]. The forward pass keeps every
// activation for the backward pass, which writes the parameter gradient in the same layout.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "pm_common.h"
#include "pm_conv.h"
#include "pm_timing.h"

namespace {

constexpr int kLayers = 15;
constexpr int kInChannels = 8;  // stored input channels (the caller fills the first few)

struct Layer {
  char name[8];
  int ca = 0, cb = 0, r = 0, up = 0, cout = 0, act = 1;
  int64_t w_off = 0, b_off = 0;
  uint16_t* wp = nullptr;   // packed forward operand
  uint16_t* wpt = nullptr;  // packed backward-data operand
  const uint16_t* xa = nullptr;
  const uint16_t* xb = nullptr;
  uint16_t* y = nullptr;    // bf16 output (hidden layers)
  float* dx = nullptr;      // gradient of the layer input (float32 [r^3][ca+cb])
};

}  // namespace

struct pm_wsnet {
  pm_wsnet_desc d{};
  int R[4] = {0, 0, 0, 0};
  int C[4] = {0, 0, 0, 0};
  Layer L[kLayers];
  int64_t nparams = 0;
  float* params = nullptr;        // device copy of the parameter vector
  uint16_t* x0 = nullptr;         // input of the last forward pass
  uint16_t* ea[4] = {};           // first conv of each encoder level
  uint16_t* eb[4] = {};           // second conv (skip tensor; eb[3] is the bottom)
  uint16_t* pool[3] = {};         // pooled eb[l] feeding level l+1
  uint16_t* da[3] = {};
  uint16_t* db[3] = {};
  uint16_t* dz = nullptr;         // output gradient of the layer being differentiated
  float* g = nullptr;             // gradient of an upsampled tensor (coarse)
  float* partial = nullptr;        // backward-weight partial sums
  int64_t partial_floats = 0;
  float* fscratch = nullptr;       // split-K partial sums of forward / backward-data passes
  int64_t fscratch_floats = 0;
  float* dout_copy = nullptr;
  bool have_params = false;
  bool have_forward = false;
  std::vector<void*> owned;
};

namespace {

pm_status nalloc(pm_wsnet* net, void** p, size_t bytes) {
  hipError_t e = hipMalloc(p, bytes > 0 ? bytes : 16);
  if (e != hipSuccess) return pm::hip_fail("hipMalloc (warm-start network)", e);
  net->owned.push_back(*p);
  return PM_OK;
}

void net_destroy(pm_wsnet* net) {
  if (!net) return;
  for (void* p : net->owned) (void)hipFree(p);
  delete net;
}

int64_t vox(int r) { return static_cast<int64_t>(r) * r * r; }

void set_layer(Layer& l, const char* name, int ca, int cb, int r, int up, int cout, int act) {
  std::snprintf(l.name, sizeof(l.name), "%s", name);
  l.ca = ca;
  l.cb = cb;
  l.r = r;
  l.up = up;
  l.cout = cout;
  l.act = act;
}

std::string timer_name(const char* pass, const Layer& l) {
  return std::string("net.") + pass + "." + l.name;
}

pm_status layer_bwd(pm_wsnet* net, int i, float* grads, int accumulate, bool need_dx,
                    hipStream_t s) {
  Layer& l = net->L[i];
  pm_status st;
  const std::string tw = timer_name("bwd_weight", l);
  if ((st = pm::conv::wgrad(l.xa, l.ca, l.xb, l.cb, l.r, l.up, net->dz, l.cout, grads + l.w_off,
                            grads + l.b_off, accumulate, net->partial, net->partial_floats,
                            tw.c_str(), s)))
    return st;
  if (!need_dx) return PM_OK;
  const std::string td = timer_name("bwd_data", l);
  return pm::conv::fwd(net->dz, pm::conv::store_channels(l.cout), nullptr, 0, l.r, 0, l.wpt,
                       nullptr, l.ca + l.cb, 0, 0.0f, nullptr, l.dx, td.c_str(), s,
                       net->fscratch, net->fscratch_floats);
}

}  // namespace

extern "C" {

pm_status pm_wsnet_create(const pm_wsnet_desc* desc, pm_wsnet** out) {
  pm::clear_error();
  PM_REQUIRE(desc && out, "pm_wsnet_create: null pointer");
  *out = nullptr;
  const pm_wsnet_desc& d = *desc;
  PM_REQUIRE(d.res >= 8 && d.res <= 256 && d.res % 8 == 0,
             "pm_wsnet_create: res must be a multiple of 8 in [8, 256]");
  PM_REQUIRE(d.in_channels >= 1 && d.in_channels <= kInChannels,
             "pm_wsnet_create: in_channels must be in [1, 8]");
  for (int l = 0; l < 4; ++l)
    PM_REQUIRE(d.width[l] == 16 || d.width[l] == 32 || d.width[l] == 64 || d.width[l] == 128,
               "pm_wsnet_create: widths must be 16, 32, 64 or 128");
  PM_REQUIRE(d.slope >= 0.0f && d.slope < 1.0f, "pm_wsnet_create: slope must be in [0, 1)");
  auto* net = new pm_wsnet();
  net->d = d;
  for (int l = 0; l < 4; ++l) {
    net->R[l] = d.res >> l;
    net->C[l] = d.width[l];
  }
  const int* R = net->R;
  const int* C = net->C;
  Layer* L = net->L;
  set_layer(L[0], "e0a", kInChannels, 0, R[0], 0, C[0], 1);
  set_layer(L[1], "e0b", C[0], 0, R[0], 0, C[0], 1);
  set_layer(L[2], "e1a", C[0], 0, R[1], 0, C[1], 1);
  set_layer(L[3], "e1b", C[1], 0, R[1], 0, C[1], 1);
  set_layer(L[4], "e2a", C[1], 0, R[2], 0, C[2], 1);
  set_layer(L[5], "e2b", C[2], 0, R[2], 0, C[2], 1);
  set_layer(L[6], "e3a", C[2], 0, R[3], 0, C[3], 1);
  set_layer(L[7], "e3b", C[3], 0, R[3], 0, C[3], 1);
  set_layer(L[8], "d2a", C[3], C[2], R[2], 1, C[2], 1);
  set_layer(L[9], "d2b", C[2], 0, R[2], 0, C[2], 1);
  set_layer(L[10], "d1a", C[2], C[1], R[1], 1, C[1], 1);
  set_layer(L[11], "d1b", C[1], 0, R[1], 0, C[1], 1);
  set_layer(L[12], "d0a", C[1], C[0], R[0], 1, C[0], 1);
  set_layer(L[13], "d0b", C[0], 0, R[0], 0, C[0], 1);
  set_layer(L[14], "out", C[0], 0, R[0], 0, 1, 0);
  int64_t off = 0;
  for (auto& l : net->L) {
    l.w_off = off;
    off += static_cast<int64_t>(l.cout) * (l.ca + l.cb) * 27;
    l.b_off = off;
    off += l.cout;
  }
  net->nparams = off;

  pm_status st = PM_OK;
  auto fail = [&](pm_status e) {
    net_destroy(net);
    return e;
  };
  auto bf = [&](uint16_t** p, int64_t elems) {
    return nalloc(net, reinterpret_cast<void**>(p), sizeof(uint16_t) * static_cast<size_t>(elems));
  };
  auto f32 = [&](float** p, int64_t elems) {
    return nalloc(net, reinterpret_cast<void**>(p), sizeof(float) * static_cast<size_t>(elems));
  };
  if ((st = f32(&net->params, net->nparams))) return fail(st);
  for (auto& l : net->L) {
    const int cin = l.ca + l.cb;
    if ((st = bf(&l.wp, pm::conv::packed_elems(l.cout, cin, 0))) ||
        (st = bf(&l.wpt, pm::conv::packed_elems(l.cout, cin, 1))) ||
        (st = f32(&l.dx, vox(l.r) * cin)))
      return fail(st);
  }
  if ((st = bf(&net->x0, vox(R[0]) * kInChannels))) return fail(st);
  for (int l = 0; l < 4; ++l) {
    if ((st = bf(&net->ea[l], vox(R[l]) * C[l])) || (st = bf(&net->eb[l], vox(R[l]) * C[l])))
      return fail(st);
  }
  for (int l = 0; l < 3; ++l) {
    if ((st = bf(&net->pool[l], vox(R[l + 1]) * C[l])) || (st = bf(&net->da[l], vox(R[l]) * C[l])) ||
        (st = bf(&net->db[l], vox(R[l]) * C[l])))
      return fail(st);
  }
  int64_t dz_elems = 0, g_elems = 0, part = 0;
  for (auto& l : net->L) {
    dz_elems = std::max<int64_t>(dz_elems, vox(l.r) * pm::conv::store_channels(l.cout));
    if (l.up) g_elems = std::max<int64_t>(g_elems, vox(l.r / 2) * l.ca);
    part = std::max<int64_t>(part, pm::conv::wgrad_partial_floats(l.ca + l.cb, l.cout, l.r));
    net->fscratch_floats = std::max<int64_t>(
        net->fscratch_floats,
        std::max(pm::conv::fwd_scratch_floats(l.ca + l.cb, l.cout, l.r),
                 pm::conv::fwd_scratch_floats(pm::conv::store_channels(l.cout), l.ca + l.cb, l.r)));
  }
  net->partial_floats = part;
  if ((st = f32(&net->fscratch, std::max<int64_t>(net->fscratch_floats, 4)))) return fail(st);
  if ((st = bf(&net->dz, dz_elems)) || (st = f32(&net->g, g_elems)) ||
      (st = f32(&net->partial, part)) || (st = f32(&net->dout_copy, vox(R[0]))))
    return fail(st);
  // wiring of inputs and outputs
  L[0].xa = net->x0;
  L[0].y = net->ea[0];
  L[1].xa = net->ea[0];
  L[1].y = net->eb[0];
  for (int l = 1; l < 4; ++l) {
    L[2 * l].xa = net->pool[l - 1];
    L[2 * l].y = net->ea[l];
    L[2 * l + 1].xa = net->ea[l];
    L[2 * l + 1].y = net->eb[l];
  }
  // decoder level l: layers 8 + 2*(2-l) and 9 + 2*(2-l)
  for (int l = 2; l >= 0; --l) {
    Layer& a = L[8 + 2 * (2 - l)];
    Layer& b = L[9 + 2 * (2 - l)];
    a.xa = (l == 2) ? net->eb[3] : net->db[l + 1];
    a.xb = net->eb[l];
    a.y = net->da[l];
    b.xa = net->da[l];
    b.y = net->db[l];
  }
  L[14].xa = net->db[0];
  L[14].y = nullptr;
  *out = net;
  return PM_OK;
}

pm_status pm_wsnet_destroy(pm_wsnet* net) {
  net_destroy(net);
  return PM_OK;
}

pm_status pm_wsnet_num_params(const pm_wsnet* net, int64_t* out) {
  PM_REQUIRE(net && out, "pm_wsnet_num_params: null pointer");
  *out = net->nparams;
  return PM_OK;
}

pm_status pm_wsnet_layer_info(const pm_wsnet* net, int32_t index, pm_wsnet_layer* out) {
  PM_REQUIRE(net && out, "pm_wsnet_layer_info: null pointer");
  PM_REQUIRE(index >= 0 && index < kLayers, "pm_wsnet_layer_info: index out of range");
  const Layer& l = net->L[index];
  std::memset(out, 0, sizeof(*out));
  std::snprintf(out->name, sizeof(out->name), "%s", l.name);
  out->cin_a = l.ca;
  out->cin_b = l.cb;
  out->cout = l.cout;
  out->res = l.r;
  out->upsample_a = l.up;
  out->act = l.act;
  out->w_offset = l.w_off;
  out->b_offset = l.b_off;
  return PM_OK;
}

pm_status pm_wsnet_set_params(pm_wsnet* net, const float* params, hipStream_t s) {
  pm::clear_error();
  PM_REQUIRE(net && params, "pm_wsnet_set_params: null pointer");
  PM_TRY_HIP(hipMemcpyAsync(net->params, params, sizeof(float) * net->nparams,
                            hipMemcpyDeviceToDevice, s));
  pm_status st;
  for (auto& l : net->L) {
    const int cin = l.ca + l.cb;
    if ((st = pm::conv::pack(net->params + l.w_off, l.cout, cin, 0, l.wp, s)) ||
        (st = pm::conv::pack(net->params + l.w_off, l.cout, cin, 1, l.wpt, s)))
      return st;
  }
  net->have_params = true;
  net->have_forward = false;
  return PM_OK;
}

pm_status pm_wsnet_forward(pm_wsnet* net, const uint16_t* input, float* out, hipStream_t s) {
  pm::clear_error();
  PM_REQUIRE(net && input && out, "pm_wsnet_forward: null pointer");
  PM_REQUIRE(net->have_params, "pm_wsnet_forward: parameters not set");
  PM_REQUIRE((reinterpret_cast<uintptr_t>(input) & 15u) == 0,
             "pm_wsnet_forward: input must be 16-byte aligned");
  PM_TRY_HIP(hipMemcpyAsync(net->x0, input, sizeof(uint16_t) * vox(net->R[0]) * kInChannels,
                            hipMemcpyDeviceToDevice, s));
  pm_status st;
  const float slope = net->d.slope;
  auto conv = [&](int i, float* y32) {
    const Layer& l = net->L[i];
    const std::string t = timer_name("fwd", l);
    return pm::conv::fwd(l.xa, l.ca, l.xb, l.cb, l.r, l.up, l.wp, net->params + l.b_off, l.cout,
                         l.act, slope, l.y, y32, t.c_str(), s, net->fscratch,
                         net->fscratch_floats);
  };
  for (int lv = 0; lv < 4; ++lv) {
    if ((st = conv(2 * lv, nullptr)) || (st = conv(2 * lv + 1, nullptr))) return st;
    if (lv < 3 && (st = pm::conv::pool2(net->eb[lv], net->R[lv], net->C[lv], net->pool[lv], s)))
      return st;
  }
  for (int i = 8; i < 14; ++i)
    if ((st = conv(i, nullptr))) return st;
  if ((st = conv(14, out))) return st;
  net->have_forward = true;
  return PM_OK;
}

pm_status pm_wsnet_backward(pm_wsnet* net, const float* dout, float* grads, int32_t accumulate,
                            hipStream_t s) {
  pm::clear_error();
  PM_REQUIRE(net && dout && grads, "pm_wsnet_backward: null pointer");
  PM_REQUIRE(net->have_forward, "pm_wsnet_backward: no forward pass to differentiate");
  const float slope = net->d.slope;
  const int* R = net->R;
  const int* C = net->C;
  Layer* L = net->L;
  pm_status st;
  auto act = [&](const float* g1, int stride, const float* g2, const uint16_t* y, int r, int c) {
    return pm::conv::act_bwd(g1, stride, g2, y, slope, r, c, pm::conv::store_channels(c), net->dz,
                             s);
  };
  // output layer (linear): dz = dout
  if ((st = act(dout, 1, nullptr, nullptr, R[0], 1))) return st;
  if ((st = layer_bwd(net, 14, grads, accumulate, true, s))) return st;
  // decoder, finest level first: d0b, d0a, d1b, d1a, d2b, d2a
  const float* g_next = L[14].dx;  // gradient of db[0]
  int g_stride = C[0];
  for (int l = 0; l <= 2; ++l) {
    const int ib = 13 - 2 * l, ia = 12 - 2 * l;
    if ((st = act(g_next, g_stride, nullptr, net->db[l], R[l], C[l]))) return st;
    if ((st = layer_bwd(net, ib, grads, accumulate, true, s))) return st;
    if ((st = act(L[ib].dx, C[l], nullptr, net->da[l], R[l], C[l]))) return st;
    if ((st = layer_bwd(net, ia, grads, accumulate, true, s))) return st;
    // the upsampled part of d_la's input is the output of the coarser decoder level (or the
    // bottom): sum the gradient over the 8 children
    if ((st = pm::conv::upsum(L[ia].dx, R[l], L[ia].ca + L[ia].cb, L[ia].ca, net->g, s)))
      return st;
    g_next = net->g;
    g_stride = L[ia].ca;
  }
  // bottom: e3b (gradient from d2a's upsampled input), e3a
  if ((st = act(g_next, g_stride, nullptr, net->eb[3], R[3], C[3]))) return st;
  if ((st = layer_bwd(net, 7, grads, accumulate, true, s))) return st;
  if ((st = act(L[7].dx, C[3], nullptr, net->ea[3], R[3], C[3]))) return st;
  if ((st = layer_bwd(net, 6, grads, accumulate, true, s))) return st;
  // encoder levels 2, 1, 0: e_lb receives the skip gradient (second part of the decoder
  // input gradient) plus the pooled path's gradient; e_la then the usual chain
  for (int l = 2; l >= 0; --l) {
    const Layer& dec = L[8 + 2 * (2 - l)];          // d_la, whose xb is eb[l]
    const float* skip = dec.dx + dec.ca;            // channels [ca, ca+cb) of its input grad
    const float* pooled = L[2 * (l + 1)].dx;        // gradient of pool[l] (e_(l+1)a input)
    if ((st = act(skip, dec.ca + dec.cb, pooled, net->eb[l], R[l], C[l]))) return st;
    if ((st = layer_bwd(net, 2 * l + 1, grads, accumulate, true, s))) return st;
    if ((st = act(L[2 * l + 1].dx, C[l], nullptr, net->ea[l], R[l], C[l]))) return st;
    if ((st = layer_bwd(net, 2 * l, grads, accumulate, l > 0, s))) return st;
  }
  return PM_OK;
}

}  // extern "C"
