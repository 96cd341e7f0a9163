// oracle_denoise.cpp - FP64 CPU restatement of the denoiser network and its
// gradients.  TEST INFRASTRUCTURE ONLY (see oracle_common.hpp).
//
// The denoiser has no reference counterpart (the reference filters with a
// joint bilateral kernel, Render.cu:353-460); this file is the oracle for the
// library's own U-Net: 3x3 'same' convolutions (NHWC, OHWI weights), ReLU,
// 2x2 average pooling, nearest upsampling with skip concatenation, residual
// prediction of log(1+radiance), L1 loss.  Gradients are by hand-written
// reverse mode.  With emulate = 1 every tensor the GPU stores in BF16 (inputs,
// activations, weights, activation gradients) is rounded to BF16 at the same
// point, so that only the accumulation order differs.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

double bf(double x) {  // round to the nearest BF16 (ties to even) through float
  float f = (float)x;
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return x;
  u += 0x7fffu + ((u >> 16) & 1u);
  u &= 0xffff0000u;
  std::memcpy(&f, &u, 4);
  return f;
}

struct T {
  int b = 0, h = 0, w = 0, c = 0;
  std::vector<double> v;
  void init(int B, int H, int W, int C) {
    b = B; h = H; w = W; c = C;
    v.assign((size_t)B * H * W * C, 0.0);
  }
  double &at(int n, int y, int x, int k) { return v[(((size_t)n * h + y) * w + x) * c + k]; }
  double at(int n, int y, int x, int k) const { return v[(((size_t)n * h + y) * w + x) * c + k]; }
};

void round16(T &t, int emulate) {
  if (emulate)
    for (double &x : t.v) x = bf(x);
}

void convFwd(const T &x, const double *W, const double *bias, int cout, T &y) {
  y.init(x.b, x.h, x.w, cout);
  for (int n = 0; n < x.b; n++)
    for (int yy = 0; yy < x.h; yy++)
      for (int xx = 0; xx < x.w; xx++)
        for (int co = 0; co < cout; co++) {
          double s = bias ? bias[co] : 0.0;
          for (int ky = 0; ky < 3; ky++) {
            int sy = yy + ky - 1;
            if (sy < 0 || sy >= x.h) continue;
            for (int kx = 0; kx < 3; kx++) {
              int sx = xx + kx - 1;
              if (sx < 0 || sx >= x.w) continue;
              const double *wr = W + (((size_t)co * 3 + ky) * 3 + kx) * x.c;
              const double *xr = &x.v[(((size_t)n * x.h + sy) * x.w + sx) * x.c];
              for (int ci = 0; ci < x.c; ci++) s += wr[ci] * xr[ci];
            }
          }
          y.at(n, yy, xx, co) = s;
        }
}

void convBwdW(const T &dy, const T &x, double *dW, double *db) {
  for (int n = 0; n < x.b; n++)
    for (int yy = 0; yy < x.h; yy++)
      for (int xx = 0; xx < x.w; xx++)
        for (int co = 0; co < dy.c; co++) {
          const double g = dy.at(n, yy, xx, co);
          if (db) db[co] += g;
          if (g == 0.0) continue;
          for (int ky = 0; ky < 3; ky++) {
            int sy = yy + ky - 1;
            if (sy < 0 || sy >= x.h) continue;
            for (int kx = 0; kx < 3; kx++) {
              int sx = xx + kx - 1;
              if (sx < 0 || sx >= x.w) continue;
              double *wr = dW + (((size_t)co * 3 + ky) * 3 + kx) * x.c;
              const double *xr = &x.v[(((size_t)n * x.h + sy) * x.w + sx) * x.c];
              for (int ci = 0; ci < x.c; ci++) wr[ci] += g * xr[ci];
            }
          }
        }
}

void convBwdX(const T &dy, const double *W, int cin, T &dx) {
  dx.init(dy.b, dy.h, dy.w, cin);
  for (int n = 0; n < dy.b; n++)
    for (int yy = 0; yy < dy.h; yy++)
      for (int xx = 0; xx < dy.w; xx++)
        for (int co = 0; co < dy.c; co++) {
          const double g = dy.at(n, yy, xx, co);
          if (g == 0.0) continue;
          for (int ky = 0; ky < 3; ky++) {
            int sy = yy + ky - 1;
            if (sy < 0 || sy >= dy.h) continue;
            for (int kx = 0; kx < 3; kx++) {
              int sx = xx + kx - 1;
              if (sx < 0 || sx >= dy.w) continue;
              const double *wr = W + (((size_t)co * 3 + ky) * 3 + kx) * cin;
              double *xr = &dx.v[(((size_t)n * dy.h + sy) * dy.w + sx) * cin];
              for (int ci = 0; ci < cin; ci++) xr[ci] += g * wr[ci];
            }
          }
        }
}

const int kLevels = 5;
const int kChan[kLevels] = {32, 64, 64, 96, 128};
const int kIn = 32, kOut = 32, kLayers = 19;

struct Net {
  int cin[kLayers], cout[kLayers];
  size_t wOff[kLayers], bOff[kLayers], n = 0;
  Net() {
    int k = 0, prev = kIn;
    for (int l = 0; l < kLevels; l++) {
      cin[k] = prev; cout[k++] = kChan[l];
      cin[k] = kChan[l]; cout[k++] = kChan[l];
      prev = kChan[l];
    }
    for (int l = kLevels - 2; l >= 0; l--) {
      cin[k] = prev + kChan[l]; cout[k++] = kChan[l];
      cin[k] = kChan[l]; cout[k++] = kChan[l];
      prev = kChan[l];
    }
    cin[k] = kChan[0]; cout[k++] = kOut;
    for (int l = 0; l < kLayers; l++) {
      wOff[l] = n;
      n += (size_t)cout[l] * 9 * cin[l];
      bOff[l] = n;
      n += cout[l];
    }
  }
};

void pool(const T &x, T &y) {
  y.init(x.b, x.h / 2, x.w / 2, x.c);
  for (int n = 0; n < x.b; n++)
    for (int yy = 0; yy < y.h; yy++)
      for (int xx = 0; xx < y.w; xx++)
        for (int c = 0; c < x.c; c++)
          y.at(n, yy, xx, c) = 0.25 * (x.at(n, 2 * yy, 2 * xx, c) + x.at(n, 2 * yy, 2 * xx + 1, c) +
                                       x.at(n, 2 * yy + 1, 2 * xx, c) + x.at(n, 2 * yy + 1, 2 * xx + 1, c));
}

void upcat(const T &low, const T &skip, T &out) {
  out.init(skip.b, skip.h, skip.w, low.c + skip.c);
  for (int n = 0; n < skip.b; n++)
    for (int y = 0; y < skip.h; y++)
      for (int x = 0; x < skip.w; x++) {
        for (int c = 0; c < low.c; c++) out.at(n, y, x, c) = low.at(n, y / 2, x / 2, c);
        for (int c = 0; c < skip.c; c++) out.at(n, y, x, low.c + c) = skip.at(n, y, x, c);
      }
}

void relu(T &t) {
  for (double &x : t.v) x = x > 0 ? x : 0;
}
void mask(T &g, const T &y) {
  for (size_t i = 0; i < g.v.size(); i++)
    if (!(y.v[i] > 0)) g.v[i] = 0;
}

}  // namespace

extern "C" {

// Single convolution probes (all buffers double; w OHWI).
void orc_conv3x3_fwd(const double *x, const double *w, const double *bias, double *y, int B, int H, int W, int Cin,
                     int Cout) {
  T X, Y;
  X.init(B, H, W, Cin);
  std::memcpy(X.v.data(), x, X.v.size() * 8);
  convFwd(X, w, bias, Cout, Y);
  std::memcpy(y, Y.v.data(), Y.v.size() * 8);
}
void orc_conv3x3_bwd_w(const double *dy, const double *x, double *dw, int B, int H, int W, int Cin, int Cout) {
  T X, DY;
  X.init(B, H, W, Cin);
  DY.init(B, H, W, Cout);
  std::memcpy(X.v.data(), x, X.v.size() * 8);
  std::memcpy(DY.v.data(), dy, DY.v.size() * 8);
  std::memset(dw, 0, (size_t)Cout * 9 * Cin * 8);
  convBwdW(DY, X, dw, nullptr);
}
void orc_conv3x3_bwd_x(const double *dy, const double *w, double *dx, int B, int H, int W, int Cin, int Cout) {
  T DY, DX;
  DY.init(B, H, W, Cout);
  std::memcpy(DY.v.data(), dy, DY.v.size() * 8);
  convBwdX(DY, w, Cin, DX);
  std::memcpy(dx, DX.v.data(), DX.v.size() * 8);
}

int64_t orc_unet_num_params() { return (int64_t)Net().n; }

// Loss and gradient of the whole network for one batch; params packed as the
// library packs them (per layer: W[Cout][3][3][Cin] then bias[Cout]).
// in10: B*h*w*10, target: B*h*w*3; h, w multiples of 16.
double orc_unet_grad(const float *params, const float *in10, const float *target, int B, int h, int w, int emulate,
                     double *grads, float *predOut) {
  Net net;
  std::vector<double> P(net.n), G(net.n, 0.0);
  for (size_t i = 0; i < net.n; i++) P[i] = params[i];
  for (int l = 0; l < kLayers; l++) {  // the GPU computes with BF16 weights
    size_t nw = (size_t)net.cout[l] * 9 * net.cin[l];
    if (emulate)
      for (size_t i = 0; i < nw; i++) P[net.wOff[l] + i] = bf(P[net.wOff[l] + i]);
  }
  auto Wp = [&](int l) { return P.data() + net.wOff[l]; };
  auto Bp = [&](int l) { return P.data() + net.bOff[l]; };
  // input
  T x0, logRad;
  x0.init(B, h, w, kIn);
  logRad.init(B, h, w, 3);
  for (int n = 0; n < B; n++)
    for (int y = 0; y < h; y++)
      for (int x = 0; x < w; x++) {
        const float *s = in10 + (((size_t)n * h + y) * w + x) * 10;
        for (int c = 0; c < 3; c++) {
          double v = emulate ? (double)log1pf(std::fmax(s[c], 0.f)) : std::log1p(std::fmax((double)s[c], 0.0));
          x0.at(n, y, x, c) = v;
          logRad.at(n, y, x, c) = v;
        }
        for (int c = 3; c < 9; c++) x0.at(n, y, x, c) = s[c];
        x0.at(n, y, x, 9) = emulate ? (double)log1pf(std::fmax(s[9], 0.f)) : std::log1p(std::fmax((double)s[9], 0.0));
      }
  round16(x0, emulate);
  // forward
  T e[kLevels][2], pl[kLevels - 1], cat[kLevels - 1], d[kLevels - 1][2], pred;
  int layer = 0;
  const T *cur = &x0;
  for (int l = 0; l < kLevels; l++) {
    if (l > 0) {
      pool(e[l - 1][1], pl[l - 1]);
      round16(pl[l - 1], emulate);
      cur = &pl[l - 1];
    }
    convFwd(*cur, Wp(layer), Bp(layer), net.cout[layer], e[l][0]);
    relu(e[l][0]);
    round16(e[l][0], emulate);
    layer++;
    convFwd(e[l][0], Wp(layer), Bp(layer), net.cout[layer], e[l][1]);
    relu(e[l][1]);
    round16(e[l][1], emulate);
    layer++;
    cur = &e[l][1];
  }
  for (int l = kLevels - 2; l >= 0; l--) {
    upcat(*cur, e[l][1], cat[l]);
    convFwd(cat[l], Wp(layer), Bp(layer), net.cout[layer], d[l][0]);
    relu(d[l][0]);
    round16(d[l][0], emulate);
    layer++;
    convFwd(d[l][0], Wp(layer), Bp(layer), net.cout[layer], d[l][1]);
    relu(d[l][1]);
    round16(d[l][1], emulate);
    layer++;
    cur = &d[l][1];
  }
  convFwd(*cur, Wp(layer), Bp(layer), net.cout[layer], pred);
  for (int n = 0; n < B; n++)
    for (int y = 0; y < h; y++)
      for (int x = 0; x < w; x++)
        for (int c = 0; c < 3; c++) pred.at(n, y, x, c) += logRad.at(n, y, x, c);
  // loss
  const double inv = 1.0 / ((double)B * h * w * 3);
  T dpred;
  dpred.init(B, h, w, kOut);
  double loss = 0.0;
  for (int n = 0; n < B; n++)
    for (int y = 0; y < h; y++)
      for (int x = 0; x < w; x++)
        for (int c = 0; c < 3; c++) {
          const float tt = target[(((size_t)n * h + y) * w + x) * 3 + c];
          double t = emulate ? (double)log1pf(std::fmax(tt, 0.f)) : std::log1p(std::fmax((double)tt, 0.0));
          double df = pred.at(n, y, x, c) - t;
          loss += std::fabs(df) * inv;
          dpred.at(n, y, x, c) = df > 0 ? inv : (df < 0 ? -inv : 0.0);
        }
  round16(dpred, emulate);
  if (predOut)
    for (size_t i = 0; i < pred.v.size(); i++) predOut[i] = (float)pred.v[i];
  // backward
  auto Gw = [&](int l) { return G.data() + net.wOff[l]; };
  auto Gb = [&](int l) { return G.data() + net.bOff[l]; };
  T gd[kLevels - 1][2], gcat[kLevels - 1], ge[kLevels][2], gpl[kLevels - 1];
  layer = kLayers - 1;
  convBwdW(dpred, d[0][1], Gw(layer), Gb(layer));
  convBwdX(dpred, Wp(layer), net.cin[layer], gd[0][1]);
  round16(gd[0][1], emulate);
  layer--;
  for (int l = 0; l <= kLevels - 2; l++) {
    mask(gd[l][1], d[l][1]);
    convBwdW(gd[l][1], d[l][0], Gw(layer), Gb(layer));
    convBwdX(gd[l][1], Wp(layer), net.cin[layer], gd[l][0]);
    round16(gd[l][0], emulate);
    layer--;
    mask(gd[l][0], d[l][0]);
    convBwdW(gd[l][0], cat[l], Gw(layer), Gb(layer));
    convBwdX(gd[l][0], Wp(layer), net.cin[layer], gcat[l]);
    round16(gcat[l], emulate);
    layer--;
    T &low = l == kLevels - 2 ? ge[kLevels - 1][1] : gd[l + 1][1];
    const T &lowFwd = l == kLevels - 2 ? e[kLevels - 1][1] : d[l + 1][1];
    low.init(B, lowFwd.h, lowFwd.w, lowFwd.c);
    ge[l][1].init(B, cat[l].h, cat[l].w, kChan[l]);
    for (int n = 0; n < B; n++)
      for (int y = 0; y < cat[l].h; y++)
        for (int x = 0; x < cat[l].w; x++) {
          for (int c = 0; c < low.c; c++) low.at(n, y / 2, x / 2, c) += gcat[l].at(n, y, x, c);
          for (int c = 0; c < kChan[l]; c++) ge[l][1].at(n, y, x, c) = gcat[l].at(n, y, x, low.c + c);
        }
    round16(low, emulate);
  }
  for (int l = kLevels - 1; l >= 0; l--) {
    mask(ge[l][1], e[l][1]);
    convBwdW(ge[l][1], e[l][0], Gw(layer), Gb(layer));
    convBwdX(ge[l][1], Wp(layer), net.cin[layer], ge[l][0]);
    round16(ge[l][0], emulate);
    layer--;
    mask(ge[l][0], e[l][0]);
    const T &in = l == 0 ? x0 : pl[l - 1];
    convBwdW(ge[l][0], in, Gw(layer), Gb(layer));
    if (l > 0) {
      convBwdX(ge[l][0], Wp(layer), net.cin[layer], gpl[l - 1]);
      round16(gpl[l - 1], emulate);
      T &g = ge[l - 1][1];
      for (int n = 0; n < B; n++)
        for (int y = 0; y < g.h; y++)
          for (int x = 0; x < g.w; x++)
            for (int c = 0; c < g.c; c++) g.at(n, y, x, c) += 0.25 * gpl[l - 1].at(n, y / 2, x / 2, c);
      round16(g, emulate);
    }
    layer--;
  }
  for (size_t i = 0; i < net.n; i++) grads[i] = G[i];
  return loss;
}

}  // extern "C"
