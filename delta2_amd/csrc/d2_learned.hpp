// d2_learned.hpp -- device code of the learned pair gate (included by
// d2_kernels.hip inside namespace d2).
//
// A 24 -> 64 -> 64 -> 2 MLP (ReLU, bf16 weights, FP32 accumulation) reads a
// pair's relative pose and motion and proposes
//   y0: the pair's conservative-advancement step scale (the TOI -> step rule's
//       scale; 1 = the pair does not limit the step),
//   y1: a lower bound on the pair's time of impact as a fraction of the CCD
//       interval (labels are 1.25 for pairs with no CCD contact at all).
// A pair proposed as contact-free (y1 and y0 above their gates) is checked by
// a sound certificate instead of being traversed: its bodies' vertex sets are
// separated, along axes found by support-point refinement, by more than the
// search distance plus every rounding and float-barycentric margin, in every
// frame the reference's sub-tests measure in.  Certified pairs provably give
// the plain path's result (no contact, scale 1, min_time = max_time); every
// other pair runs the plain traversal.

constexpr int MLP_IN = 24, MLP_H = 64, MLP_OUT = 2;
constexpr int OFF_W1 = 0;                            // W1[24][64], row = input
constexpr int OFF_B1 = OFF_W1 + MLP_IN * MLP_H;      // b1[64]
constexpr int OFF_W2 = OFF_B1 + MLP_H;               // W2[64][64]
constexpr int OFF_B2 = OFF_W2 + MLP_H * MLP_H;       // b2[64]
constexpr int OFF_W3 = OFF_B2 + MLP_H;               // W3[64][2]
constexpr int OFF_B3 = OFF_W3 + MLP_H * MLP_OUT;     // b3[2]
constexpr int MLP_PARAMS = OFF_B3 + MLP_OUT;         // 5890
constexpr int MLP_ACT = MLP_IN + 2 * MLP_H;          // stored activations per sample
constexpr int MLP_DELTA = 2 * MLP_H + MLP_OUT;       // stored deltas per sample

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

// float -> bf16, round to nearest even (inputs are finite).
__device__ __forceinline__ short f2bf(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}
__device__ __forceinline__ float bf2f(short h) { return __uint_as_float(((unsigned)(unsigned short)h) << 16); }

// ------------------------------------------------------------------ features
// Per pair, in body a's frame at the start of the interval, lengths in units
// of Lc = r_a + r_b (bounding radii), times in units of the interval:
//   0-2 centre offset, 3-6 relative rotation (w >= 0), 7-9 relative velocity,
//   10-12 / 13-15 angular velocities of a / b, 16-17 radii, 18 search
//   distance, 19 centre distance, 20 gap between the bounding spheres,
//   21 closing speed, 22 rotational sweep (|w_a| r_a + |w_b| r_b), 23 |v_rel|.
__device__ __forceinline__ void pair_features(int p, const int32_t* pairs, const double* bodies,
                                              const int32_t* body_tree, const float* tree_radius,
                                              const double* max_time, float* x) {
  {
    const int a = pairs[2 * p], b = pairs[2 * p + 1];
    const double* ra = bodies + 40 * (size_t)a;
    const double* rb = bodies + 40 * (size_t)b;
    const body_state sa = load_state(ra), sb = load_state(rb);
    const double mt = max_time ? max_time[p] : ((rb[39] < ra[39]) ? rb[39] : ra[39]);
    const double rA = tree_radius[body_tree[a]], rB = tree_radius[body_tree[b]];
    const double Lc = rA + rB > 1e-12 ? rA + rB : 1e-12;
    double R[9], Iw[9];
    qrot(sa.q, R);
    auto to_a = [&](v3<double> v) {  // R^T v
      return v3<double>{(R[0] * v.x + R[3] * v.y) + R[6] * v.z, (R[1] * v.x + R[4] * v.y) + R[7] * v.z,
                        (R[2] * v.x + R[5] * v.y) + R[8] * v.z};
    };
    world_inv_inertia(sa.q, ra + 28, Iw);
    const v3<double> wa = mul33v(Iw, sa.L);
    world_inv_inertia(sb.q, rb + 28, Iw);
    const v3<double> wb = mul33v(Iw, sb.L);
    const v3<double> dw = sub(sb.t, sa.t);
    const v3<double> d = to_a(dw);
    quat qc{sa.q.w, -sa.q.x, -sa.q.y, -sa.q.z};
    quat qr = qmul(qc, sb.q);
    if (qr.w < 0.0) qr = quat{-qr.w, -qr.x, -qr.y, -qr.z};
    const v3<double> vr = to_a(sub(sb.v, sa.v));
    const v3<double> wa_a = to_a(wa), wb_a = to_a(wb);
    const double dist = vlen(d), vlenr = vlen(vr);
    const double closing = dist > 0.0 ? -dot(d, vr) / dist : 0.0;
    const double il = 1.0 / Lc, vt = mt / Lc;
    x[0] = (float)(d.x * il), x[1] = (float)(d.y * il), x[2] = (float)(d.z * il);
    x[3] = (float)qr.w, x[4] = (float)qr.x, x[5] = (float)qr.y, x[6] = (float)qr.z;
    x[7] = (float)(vr.x * vt), x[8] = (float)(vr.y * vt), x[9] = (float)(vr.z * vt);
    x[10] = (float)(wa_a.x * mt), x[11] = (float)(wa_a.y * mt), x[12] = (float)(wa_a.z * mt);
    x[13] = (float)(wb_a.x * mt), x[14] = (float)(wb_a.y * mt), x[15] = (float)(wb_a.z * mt);
    x[16] = (float)(rA * il), x[17] = (float)(rB * il);
    x[18] = (float)((double)(float)(ra[37] + rb[37]) * il);
    x[19] = (float)(dist * il);
    x[20] = (float)((dist - rA - rB) * il);
    x[21] = (float)(closing * vt);
    x[22] = (float)((vlen(wa) * rA + vlen(wb) * rB) * vt);
    x[23] = (float)(vlenr * vt);
  }
}
__global__ void k_pair_features(int n_pairs, const int32_t* pairs, const double* bodies, const int32_t* body_tree,
                                const float* tree_radius, const double* max_time, float* X) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n_pairs; p += gridDim.x * blockDim.x)
    pair_features(p, pairs, bodies, body_tree, tree_radius, max_time, X + (size_t)MLP_IN * p);
}

// --------------------------------------------------------- MFMA inference
// One wavefront per tile of 16 pairs, v_mfma_f32_16x16x16_bf16 (FP32
// accumulation).  Operand layout: lane l supplies A[l%16][4(l/16)+i] and
// B[4(l/16)+i][l%16], i = 0..3, and receives D[4(l/16)+i][l%16].
// Activations go through LDS (bf16) between layers.  wb: bf16 parameters.
// FEAT: the inputs are the pairs' features, computed here (lanes 0-15 of the
// wavefront, one pair each, into LDS) instead of read from X.
struct FeatIn {
  const int32_t* pairs;
  const double* bodies;
  const int32_t* body_tree;
  const float* tree_radius;
  const double* max_time;
};
template <bool FEAT>
__global__ void __launch_bounds__(256) k_mlp_infer(int n, const float* X, const short* wb, float* Y, FeatIn F) {
  __shared__ float s_x[FEAT ? 4 : 1][16 * MLP_IN];
  __shared__ short w1[32 * MLP_H];  // K padded to 32
  __shared__ short w2[MLP_H * MLP_H];
  __shared__ short w3[MLP_H * 16];  // N padded to 16
  __shared__ float bias[2 * MLP_H + MLP_OUT];
  __shared__ short act[4][16 * MLP_H];
  for (int i = threadIdx.x; i < 32 * MLP_H; i += blockDim.x) {
    const int k = i / MLP_H;
    w1[i] = k < MLP_IN ? wb[OFF_W1 + i] : (short)0;
  }
  for (int i = threadIdx.x; i < MLP_H * MLP_H; i += blockDim.x) w2[i] = wb[OFF_W2 + i];
  for (int i = threadIdx.x; i < MLP_H * 16; i += blockDim.x) {
    const int k = i / 16, c = i % 16;
    w3[i] = c < MLP_OUT ? wb[OFF_W3 + k * MLP_OUT + c] : (short)0;
  }
  for (int i = threadIdx.x; i < MLP_H; i += blockDim.x) bias[i] = bf2f(wb[OFF_B1 + i]), bias[MLP_H + i] = bf2f(wb[OFF_B2 + i]);
  if (threadIdx.x < MLP_OUT) bias[2 * MLP_H + threadIdx.x] = bf2f(wb[OFF_B3 + threadIdx.x]);
  __syncthreads();
  const int lane = lane_id(), wid = threadIdx.x >> 6;
  const int r = lane & 15, kq = lane >> 4;
  short* A = act[wid];
  const int n_tiles = (n + 15) / 16;
  for (int base = blockIdx.x * 4; base < n_tiles; base += gridDim.x * 4) {  // same trip count in every wave
    const int tile = base + wid;
    const int row = tile * 16 + r;
    const bool valid_row = tile < n_tiles && row < n;
    floatx4 acc[4];
    if (FEAT) {
      if (kq == 0 && valid_row)
        pair_features(row, F.pairs, F.bodies, F.body_tree, F.tree_radius, F.max_time, s_x[wid] + MLP_IN * r);
      wave_sync_lds();
    }
    // layer 1
    for (int nt = 0; nt < 4; ++nt) acc[nt] = floatx4{0.f, 0.f, 0.f, 0.f};
    for (int ks = 0; ks < 2; ++ks) {
      bf16x4 a;
      for (int i = 0; i < 4; ++i) {
        const int k = ks * 16 + 4 * kq + i;
        const float xv = !(valid_row && k < MLP_IN) ? 0.f : FEAT ? s_x[FEAT ? wid : 0][MLP_IN * r + k]
                                                                  : X[(size_t)row * MLP_IN + k];
        a[i] = (valid_row && k < MLP_IN) ? f2bf(xv) : (short)0;
      }
      for (int nt = 0; nt < 4; ++nt) {
        bf16x4 b;
        for (int i = 0; i < 4; ++i) b[i] = w1[(ks * 16 + 4 * kq + i) * MLP_H + nt * 16 + r];
        acc[nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[nt], 0, 0, 0);
      }
    }
    for (int nt = 0; nt < 4; ++nt)
      for (int i = 0; i < 4; ++i) {
        const int m = 4 * kq + i, c = nt * 16 + r;
        const float v = acc[nt][i] + bias[c];
        A[m * MLP_H + c] = f2bf(v > 0.f ? v : 0.f);
      }
    __syncthreads();
    // layer 2
    for (int nt = 0; nt < 4; ++nt) acc[nt] = floatx4{0.f, 0.f, 0.f, 0.f};
    for (int ks = 0; ks < 4; ++ks) {
      bf16x4 a;
      for (int i = 0; i < 4; ++i) a[i] = A[r * MLP_H + ks * 16 + 4 * kq + i];
      for (int nt = 0; nt < 4; ++nt) {
        bf16x4 b;
        for (int i = 0; i < 4; ++i) b[i] = w2[(ks * 16 + 4 * kq + i) * MLP_H + nt * 16 + r];
        acc[nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[nt], 0, 0, 0);
      }
    }
    __syncthreads();
    for (int nt = 0; nt < 4; ++nt)
      for (int i = 0; i < 4; ++i) {
        const int m = 4 * kq + i, c = nt * 16 + r;
        const float v = acc[nt][i] + bias[MLP_H + c];
        A[m * MLP_H + c] = f2bf(v > 0.f ? v : 0.f);
      }
    __syncthreads();
    // layer 3
    floatx4 out{0.f, 0.f, 0.f, 0.f};
    for (int ks = 0; ks < 4; ++ks) {
      bf16x4 a, b;
      for (int i = 0; i < 4; ++i) {
        a[i] = A[r * MLP_H + ks * 16 + 4 * kq + i];
        b[i] = w3[(ks * 16 + 4 * kq + i) * 16 + r];
      }
      out = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, out, 0, 0, 0);
    }
    if (r < MLP_OUT && tile < n_tiles)
      for (int i = 0; i < 4; ++i) {
        const int orow = tile * 16 + 4 * kq + i;
        if (orow < n) Y[(size_t)orow * MLP_OUT + r] = out[i] + bias[2 * MLP_H + r];
      }
    __syncthreads();
  }
}

// --------------------------------------------------------------- training
// Forward and backward of one minibatch, one thread per sample, FP32 with the
// bf16-rounded weights wf (the network inference evaluates).  Loss per sample:
// sum_o lw_o * huber_delta(y_o - t_o); the batch loss is its mean.
__global__ void __launch_bounds__(256) k_mlp_fwd_bwd(int nb, const int* idx, const float* X, const float* T,
                                                     const float* wf, float lw0, float lw1, float delta_h, float* act,
                                                     float* dlt, float* loss) {
  __shared__ float w[MLP_PARAMS];
  for (int i = threadIdx.x; i < MLP_PARAMS; i += blockDim.x) w[i] = wf[i];
  __syncthreads();
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nb) return;
  const int src = idx ? idx[s] : s;
  float x[MLP_IN], h1[MLP_H], h2[MLP_H];
  for (int k = 0; k < MLP_IN; ++k) x[k] = X[(size_t)src * MLP_IN + k];
  for (int j = 0; j < MLP_H; ++j) {
    float z = w[OFF_B1 + j];
    for (int k = 0; k < MLP_IN; ++k) z += x[k] * w[OFF_W1 + k * MLP_H + j];
    h1[j] = z > 0.f ? z : 0.f;
  }
  for (int j = 0; j < MLP_H; ++j) {
    float z = w[OFF_B2 + j];
    for (int k = 0; k < MLP_H; ++k) z += h1[k] * w[OFF_W2 + k * MLP_H + j];
    h2[j] = z > 0.f ? z : 0.f;
  }
  float y[MLP_OUT], d3[MLP_OUT];
  const float lw[MLP_OUT] = {lw0, lw1};
  float l = 0.f;
  for (int o = 0; o < MLP_OUT; ++o) {
    float z = w[OFF_B3 + o];
    for (int k = 0; k < MLP_H; ++k) z += h2[k] * w[OFF_W3 + k * MLP_OUT + o];
    y[o] = z;
    const float rr = y[o] - T[(size_t)src * MLP_OUT + o], ar = fabsf(rr);
    l += lw[o] * (ar <= delta_h ? 0.5f * rr * rr : delta_h * (ar - 0.5f * delta_h));
    const float g = ar <= delta_h ? rr : (rr > 0.f ? delta_h : -delta_h);
    d3[o] = lw[o] * g / (float)nb;
  }
  loss[s] = l;
  float* A = act + (size_t)s * MLP_ACT;
  float* D = dlt + (size_t)s * MLP_DELTA;
  for (int k = 0; k < MLP_IN; ++k) A[k] = x[k];
  for (int j = 0; j < MLP_H; ++j) A[MLP_IN + j] = h1[j], A[MLP_IN + MLP_H + j] = h2[j];
  float d2v[MLP_H];
  for (int j = 0; j < MLP_H; ++j) {
    float g = 0.f;
    for (int o = 0; o < MLP_OUT; ++o) g += w[OFF_W3 + j * MLP_OUT + o] * d3[o];
    d2v[j] = h2[j] > 0.f ? g : 0.f;
  }
  for (int i = 0; i < MLP_H; ++i) {
    float g = 0.f;
    for (int j = 0; j < MLP_H; ++j) g += w[OFF_W2 + i * MLP_H + j] * d2v[j];
    D[i] = h1[i] > 0.f ? g : 0.f;  // delta of layer 1
  }
  for (int j = 0; j < MLP_H; ++j) D[MLP_H + j] = d2v[j];
  for (int o = 0; o < MLP_OUT; ++o) D[2 * MLP_H + o] = d3[o];
}

// Parameter gradients: one thread per parameter, a fixed-order sum over the
// minibatch (deterministic).
__global__ void k_mlp_grad(int nb, const float* act, const float* dlt, float* grad) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= MLP_PARAMS) return;
  int ai, di;  // activation index (-1: bias) and delta index
  if (q < OFF_B1) ai = q / MLP_H, di = q % MLP_H;
  else if (q < OFF_W2) ai = -1, di = q - OFF_B1;
  else if (q < OFF_B2) ai = MLP_IN + (q - OFF_W2) / MLP_H, di = MLP_H + (q - OFF_W2) % MLP_H;
  else if (q < OFF_W3) ai = -1, di = MLP_H + (q - OFF_B2);
  else if (q < OFF_B3) ai = MLP_IN + MLP_H + (q - OFF_W3) / MLP_OUT, di = 2 * MLP_H + (q - OFF_W3) % MLP_OUT;
  else ai = -1, di = 2 * MLP_H + (q - OFF_B3);
  float g = 0.f;
  for (int s = 0; s < nb; ++s) {
    const float d = dlt[(size_t)s * MLP_DELTA + di];
    g += ai < 0 ? d : act[(size_t)s * MLP_ACT + ai] * d;
  }
  grad[q] = g;
}

// Adam on the FP32 master weights; refreshes the bf16 copy and its float view.
__global__ void k_mlp_adam(float* w, float* m, float* v, const float* g, float lr, float beta1, float beta2, float eps,
                           float c1, float c2, short* wb, float* wf) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= MLP_PARAMS) return;
  const float gq = g[q];
  const float mq = beta1 * m[q] + (1.f - beta1) * gq;
  const float vq = beta2 * v[q] + (1.f - beta2) * gq * gq;
  m[q] = mq, v[q] = vq;
  const float wq = w[q] - lr * (mq / c1) / (sqrtf(vq / c2) + eps);
  w[q] = wq;
  const short h = f2bf(wq);
  wb[q] = h;
  wf[q] = bf2f(h);
}
__global__ void k_mlp_round(const float* w, short* wb, float* wf) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= MLP_PARAMS) return;
  const short h = f2bf(w[q]);
  wb[q] = h;
  wf[q] = bf2f(h);
}

// ------------------------------------------------------------- certificate
// Lower bounds of the distances the reference's sub-tests can report, from
// separating axes.  Along a unit axis u, set P lies below set Q by
// min_Q u.q - max_P u.p, a lower bound of their distance.  Three bounds:
//   world, per time: the CCD's edge-edge sub-tests lerp world positions, so a
//     point pair's offset at time t is the lerp of its offsets at the ends and
//     min(gap at start, gap at end) along one axis bounds every t;
//   b's frame: a point of a lerps between its start and end positions in b's
//     frame (point-triangle sub-tests) while b is fixed: the gap between the
//     union of a's two poses and b;
//   a's frame: the same with the bodies swapped.
// The discrete phase needs one bound: the two bodies at the future poses.
// One block per pair; a set's vertices are transformed into LDS once; axes
// come from Gilbert's minimum-norm-point iteration on the difference set
// (every iterate's axis gives a sound bound, the best one is kept).
constexpr int CERT_MAXV = 1024;  // vertices per body held in LDS; larger bodies are not certified
__device__ __forceinline__ void cert_mat(const double* m, const double* origin, float* o) {
  for (int r = 0; r < 3; ++r) {
    o[4 * r] = (float)m[4 * r], o[4 * r + 1] = (float)m[4 * r + 1], o[4 * r + 2] = (float)m[4 * r + 2];
    o[4 * r + 3] = (float)(m[4 * r + 3] - (origin ? origin[r] : 0.0));
  }
}
// LDS copy of n body-frame vertices under the 3x4 matrix m (identity when null).
__device__ void cert_load(const float* v, int n, const float* m, float* out, float& mag) {
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    const float x = v[3 * k], y = v[3 * k + 1], z = v[3 * k + 2];
    float o0 = x, o1 = y, o2 = z;
    if (m) {
      o0 = (m[0] * x + m[1] * y) + (m[2] * z + m[3]);
      o1 = (m[4] * x + m[5] * y) + (m[6] * z + m[7]);
      o2 = (m[8] * x + m[9] * y) + (m[10] * z + m[11]);
    }
    out[3 * k] = o0, out[3 * k + 1] = o1, out[3 * k + 2] = o2;
    mag = fmaxf(mag, fmaxf(fabsf(o0), fmaxf(fabsf(o1), fabsf(o2))));
  }
}
// Block-wide extreme of u.x over sets S[0..ns) (sign +1: max, -1: min) and
// the point attaining it.
__device__ float cert_extreme(const float* const* S, const int* n, int ns, const float* u, float sign, float* pt,
                              float* red) {
  float best = -3.4e38f, bp[3] = {0.f, 0.f, 0.f};
  for (int s = 0; s < ns; ++s)
    for (int k = threadIdx.x; k < n[s]; k += blockDim.x) {
      const float* q = S[s] + 3 * k;
      const float pr = sign * ((u[0] * q[0] + u[1] * q[1]) + u[2] * q[2]);
      if (pr > best) best = pr, bp[0] = q[0], bp[1] = q[1], bp[2] = q[2];
    }
  for (int off = 32; off > 0; off >>= 1) {
    const float ob = __shfl_xor(best, off);
    const float o0 = __shfl_xor(bp[0], off), o1 = __shfl_xor(bp[1], off), o2 = __shfl_xor(bp[2], off);
    if (ob > best) best = ob, bp[0] = o0, bp[1] = o1, bp[2] = o2;
  }
  __syncthreads();
  if ((threadIdx.x & 63) == 0) {
    float* r = red + 4 * (threadIdx.x >> 6);
    r[0] = best, r[1] = bp[0], r[2] = bp[1], r[3] = bp[2];
  }
  __syncthreads();
  int w = 0;
  for (int i = 1; i < (int)(blockDim.x >> 6); ++i)
    if (red[4 * i] > red[4 * w]) w = i;
  pt[0] = red[4 * w + 1], pt[1] = red[4 * w + 2], pt[2] = red[4 * w + 3];
  return sign * red[4 * w];
}
// Separation bound of P below Q along u: (min_Q u.q - max_P u.p) / |u|, and
// the support points.
__device__ float cert_axis(const float* const* P, const int* nP, int nsP, const float* const* Q, const int* nQ, int nsQ,
                           const float* u, float* pp, float* qq, float* red) {
  const float nu = sqrtf((u[0] * u[0] + u[1] * u[1]) + u[2] * u[2]);
  const float hp = cert_extreme(P, nP, nsP, u, 1.f, pp, red);
  const float lq = cert_extreme(Q, nQ, nsQ, u, -1.f, qq, red);
  return nu > 1e-30f ? (lq - hp) / nu : -3.4e38f;
}
// Gilbert's minimum-norm-point iteration on the difference set Q - P from the
// point x0 (a member of its hull): the support point s = argmin (x.d) over
// Q - P bounds the distance by x.s / |x| at every step (kept as the best
// bound), and x moves to the point of segment [x, s] closest to the origin.
// Returns the best bound; u_out = its axis.
// With stop >= 0 the iteration ends as soon as the answer to "best > stop"
// is known: a gap above stop, or an iterate x with |x| <= stop (x lies in the
// hull of the difference set, so no axis can separate it by more than |x|).
__device__ float cert_gilbert(const float* const* P, const int* nP, int nsP, const float* const* Q, const int* nQ,
                              int nsQ, const float* x0, int iters, float* u_out, float* red, float stop = -1.f) {
  float x[3] = {x0[0], x0[1], x0[2]};
  float best = -3.4e38f;
  u_out[0] = x[0], u_out[1] = x[1], u_out[2] = x[2];
  for (int it = 0; it < iters; ++it) {
    const float xx = (x[0] * x[0] + x[1] * x[1]) + x[2] * x[2];
    if (!(xx > 1e-30f)) break;
    if (stop >= 0.f && it > 0 && xx <= stop * stop) break;
    float pp[3], qq[3];
    const float g = cert_axis(P, nP, nsP, Q, nQ, nsQ, x, pp, qq, red);
    if (g > best) best = g, u_out[0] = x[0], u_out[1] = x[1], u_out[2] = x[2];
    if (stop >= 0.f && best > stop) break;
    const float sd[3] = {qq[0] - pp[0], qq[1] - pp[1], qq[2] - pp[2]};
    const float xs = (x[0] * sd[0] + x[1] * sd[1]) + x[2] * sd[2];
    if (xx - xs <= 1e-7f * xx) break;  // x is (nearly) the minimum-norm point
    const float e[3] = {x[0] - sd[0], x[1] - sd[1], x[2] - sd[2]};
    const float ee = (e[0] * e[0] + e[1] * e[1]) + e[2] * e[2];
    float lam = ee > 0.f ? (xx - xs) / ee : 0.f;  // x + lam (s - x) is closest to the origin
    lam = fminf(fmaxf(lam, 0.f), 1.f);
    x[0] += lam * (sd[0] - x[0]), x[1] += lam * (sd[1] - x[1]), x[2] += lam * (sd[2] - x[2]);
  }
  return best;
}
// Per-time world bound: min over the two instants of the separation along one
// axis, the axis taken from each instant's own Gilbert iteration.
__device__ float cert_gap_timed(const float* P0, const float* P1, int nP, const float* Q0, const float* Q1, int nQ,
                                const float* x0, const float* x1, int iters, float* red) {
  const float* Ps[2] = {P0, P1};
  const float* Qs[2] = {Q0, Q1};
  float ax[3][3];
  for (int k = 0; k < 2; ++k) (void)cert_gilbert(Ps + k, &nP, 1, Qs + k, &nQ, 1, k ? x1 : x0, iters, ax[k], red);
  const float n0 = sqrtf((ax[0][0] * ax[0][0] + ax[0][1] * ax[0][1]) + ax[0][2] * ax[0][2]);
  const float n1 = sqrtf((ax[1][0] * ax[1][0] + ax[1][1] * ax[1][1]) + ax[1][2] * ax[1][2]);
  for (int r = 0; r < 3; ++r) ax[2][r] = (n0 > 0.f ? ax[0][r] / n0 : 0.f) + (n1 > 0.f ? ax[1][r] / n1 : 0.f);
  float best = -3.4e38f;
  for (int c = 0; c < 3; ++c) {
    float pp[3], qq[3];
    const float g0 = cert_axis(Ps, &nP, 1, Qs, &nQ, 1, ax[c], pp, qq, red);
    const float g1 = cert_axis(Ps + 1, &nP, 1, Qs + 1, &nQ, 1, ax[c], pp, qq, red);
    best = fmaxf(best, fminf(g0, g1));
  }
  return best;
}
constexpr int CERT_ITERS = 16;
constexpr int CERT_LIST = 1024;  // pairs per block at most (the grid is sized for it)

// gate: the MLP outputs (y0 >= s_gate and y1 >= t_gate), or every pair when
// null.  keep[p] = 0 for a certified pair, 1 otherwise.  counters[0..1]:
// proposed / certified pairs.  The kept pairs are the phase's roots: each
// block appends its kept pairs to soup with one atomic on *n_soup (zeroed by
// the phase's set-up; the discrete and CCD lists need no order).
__global__ void __launch_bounds__(256) k_certify(int n_pairs, const int32_t* pair_info, const int32_t* body_tree,
                                                 const double* pair_rec, const float* Y, float t_gate, float s_gate,
                                                 const int32_t* hv_off, const float* hv, const float* tree_eps_out,
                                                 int mode, int* keep, unsigned long long* counters, NP* soup,
                                                 unsigned long long* n_soup) {
  __shared__ float buf[4][3 * CERT_MAXV];
  __shared__ float red[4 * 16];
  __shared__ float mag_red[16];
  __shared__ int s_kept[CERT_LIST];
  __shared__ int s_nk;
  __shared__ unsigned long long s_base;
  const float EPSF = 5.9604645e-08f;
  if (threadIdx.x == 0) s_nk = 0;
  for (int p = blockIdx.x; p < n_pairs; p += gridDim.x) {
    const bool gated = !Y || (Y[2 * p] >= s_gate && Y[2 * p + 1] >= t_gate);
    const int a = pair_info[4 * p], b = pair_info[4 * p + 1];
    const int ta = body_tree[a], tb = body_tree[b];
    const int na = hv_off[ta + 1] - hv_off[ta], nb = hv_off[tb + 1] - hv_off[tb];
    if (!gated || na > CERT_MAXV || nb > CERT_MAXV) {
      if (threadIdx.x == 0) {
        keep[p] = 1;
        s_kept[s_nk++] = p;
        if (gated) atomicAdd(&counters[0], 1ull);
      }
      continue;
    }
    const double* pr = pair_rec + (size_t)PR_STRIDE * p;
    const float thr = (float)pr[PR_SEARCH] + tree_eps_out[ta] + tree_eps_out[tb];
    const float* va = hv + 3 * (size_t)hv_off[ta];
    const float* vb = hv + 3 * (size_t)hv_off[tb];
    float m0[12], m1[12], m2[12], m3[12], u[3];
    float mag = 0.f;
    bool ok = true;
    auto finish_mag = [&]() {  // block maximum of the coordinate magnitudes
      float m = mag;
      for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
      __syncthreads();
      if ((threadIdx.x & 63) == 0) mag_red[threadIdx.x >> 6] = m;
      __syncthreads();
      float r = 0.f;
      for (int i = 0; i < (int)(blockDim.x >> 6); ++i) r = fmaxf(r, mag_red[i]);
      return r;
    };
    if (mode == 0) {
      const double o[3] = {pr[PR_TB0 + 3], pr[PR_TB0 + 7], pr[PR_TB0 + 11]};
      cert_mat(pr + PR_TA0, o, m0), cert_mat(pr + PR_TA1, o, m1), cert_mat(pr + PR_TB0, o, m2), cert_mat(pr + PR_TB1, o, m3);
      cert_load(va, na, m0, buf[0], mag), cert_load(va, na, m1, buf[1], mag);
      cert_load(vb, nb, m2, buf[2], mag), cert_load(vb, nb, m3, buf[3], mag);
      const float M = finish_mag();
      {  // world, per instant: (a at start, b at start), (a at end, b at end)
        float x0[3], x1[3];
        for (int k = 0; k < 3; ++k) x0[k] = m2[4 * k + 3] - m0[4 * k + 3], x1[k] = m3[4 * k + 3] - m1[4 * k + 3];
        ok = cert_gap_timed(buf[0], buf[1], na, buf[2], buf[3], nb, x0, x1, CERT_ITERS, red) -
                 (1e-6f + 64.f * EPSF * M) > thr;
      }
      if (ok) {  // a's two poses in b's frame against b
        __syncthreads();
        mag = 0.f;
        cert_mat(pr + PR_MAB0, nullptr, m0), cert_mat(pr + PR_MAB1, nullptr, m1);
        cert_load(va, na, m0, buf[0], mag), cert_load(va, na, m1, buf[1], mag), cert_load(vb, nb, nullptr, buf[2], mag);
        const float M2 = finish_mag();
        for (int k = 0; k < 3; ++k) u[k] = -0.5f * (m0[4 * k + 3] + m1[4 * k + 3]);
        const float* P[2] = {buf[0], buf[1]};
        const float* Q[1] = {buf[2]};
        const int nP[2] = {na, na}, nQ[1] = {nb};
        float ax[3];
        ok = cert_gilbert(P, nP, 2, Q, nQ, 1, u, CERT_ITERS, ax, red) - (1e-6f + 64.f * EPSF * M2) > thr;
      }
      if (ok) {  // b's two poses in a's frame against a
        __syncthreads();
        mag = 0.f;
        cert_mat(pr + PR_MBA0, nullptr, m0), cert_mat(pr + PR_MBA1, nullptr, m1);
        cert_load(vb, nb, m0, buf[0], mag), cert_load(vb, nb, m1, buf[1], mag), cert_load(va, na, nullptr, buf[2], mag);
        const float M3 = finish_mag();
        for (int k = 0; k < 3; ++k) u[k] = -0.5f * (m0[4 * k + 3] + m1[4 * k + 3]);
        const float* P[2] = {buf[0], buf[1]};
        const float* Q[1] = {buf[2]};
        const int nP[2] = {nb, nb}, nQ[1] = {na};
        float ax[3];
        ok = cert_gilbert(P, nP, 2, Q, nQ, 1, u, CERT_ITERS, ax, red) - (1e-6f + 64.f * EPSF * M3) > thr;
      }
    } else {
      const double o[3] = {pr[PR_TB1 + 3], pr[PR_TB1 + 7], pr[PR_TB1 + 11]};
      cert_mat(pr + PR_TA1, o, m0), cert_mat(pr + PR_TB1, o, m1);
      cert_load(va, na, m0, buf[0], mag), cert_load(vb, nb, m1, buf[1], mag);
      const float M = finish_mag();
      for (int k = 0; k < 3; ++k) u[k] = m1[4 * k + 3] - m0[4 * k + 3];
      const float* P[1] = {buf[0]};
      const float* Q[1] = {buf[1]};
      const int nP[1] = {na}, nQ[1] = {nb};
      float ax[3];
      const float margin = 1e-6f + 64.f * EPSF * M;
      ok = cert_gilbert(P, nP, 1, Q, nQ, 1, u, CERT_ITERS, ax, red, thr + margin) - margin > thr;
    }
    if (threadIdx.x == 0) {
      keep[p] = ok ? 0 : 1;
      if (!ok) s_kept[s_nk++] = p;
      atomicAdd(&counters[0], 1ull);
      if (ok) atomicAdd(&counters[1], 1ull);
    }
    __syncthreads();  // LDS reuse by the next pair
  }
  __syncthreads();
  if (threadIdx.x == 0) s_base = s_nk ? atomicAdd(n_soup, (unsigned long long)s_nk) : 0ull;
  __syncthreads();
  for (int i = threadIdx.x; i < s_nk; i += blockDim.x)
    soup[s_base + i] = NP{(uint32_t)s_kept[i], 0u, 0u, 0u, 0ull};
}

