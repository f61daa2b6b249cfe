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
__global__ void k_pair_features(int n_pairs, const int32_t* pairs, const double* bodies, const int32_t* body_tree,
                                const float* tree_radius, const double* max_time, float* X) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n_pairs; p += gridDim.x * blockDim.x) {
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
    float* x = X + (size_t)MLP_IN * p;
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

// --------------------------------------------------------- MFMA inference
// One wavefront per tile of 16 pairs, v_mfma_f32_16x16x16_bf16 (FP32
// accumulation).  Operand layout: lane l supplies A[l%16][4(l/16)+i] and
// B[4(l/16)+i][l%16], i = 0..3, and receives D[4(l/16)+i][l%16].
// Activations go through LDS (bf16) between layers.  wb: bf16 parameters.
__global__ void __launch_bounds__(256) k_mlp_infer(int n, const float* X, const short* wb, float* Y) {
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
    // layer 1
    for (int nt = 0; nt < 4; ++nt) acc[nt] = floatx4{0.f, 0.f, 0.f, 0.f};
    for (int ks = 0; ks < 2; ++ks) {
      bf16x4 a;
      for (int i = 0; i < 4; ++i) {
        const int k = ks * 16 + 4 * kq + i;
        a[i] = (valid_row && k < MLP_IN) ? f2bf(X[(size_t)row * MLP_IN + k]) : (short)0;
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
// Lower bound of the distance between two vertex sets in one frame: the gap
// of their projections on an axis (any axis bounds the distance from below);
// the axis starts at the centre difference and is refined three times to the
// difference of the support points.  Vertices are given in body coordinates
// and mapped by 3x4 FP32 matrices (one or two poses per set: the sets of a
// swept set are the union of both poses, which holds every linear blend).
struct CertSet {
  const float* v;  // xyz per vertex
  int n;
  float m[2][12];
  int poses;
};
__device__ __forceinline__ void cert_xform(const float* m, const float* v, float* o) {
  o[0] = (m[0] * v[0] + m[1] * v[1]) + (m[2] * v[2] + m[3]);
  o[1] = (m[4] * v[0] + m[5] * v[1]) + (m[6] * v[2] + m[7]);
  o[2] = (m[8] * v[0] + m[9] * v[1]) + (m[10] * v[2] + m[11]);
}
// support of set S along u (sign = +1: max, -1: min); returns the extreme
// projection and the point; also the largest coordinate magnitude seen.
__device__ float cert_support(const CertSet& S, const float* u, float sign, float* pt, float& mag) {
  const int lane = lane_id();
  float best = -3.4e38f, bp[3] = {0.f, 0.f, 0.f};
  for (int k = lane; k < S.n * S.poses; k += 64) {
    float o[3];
    cert_xform(S.m[k / S.n], S.v + 3 * (k % S.n), o);
    const float pr = sign * ((u[0] * o[0] + u[1] * o[1]) + u[2] * o[2]);
    mag = fmaxf(mag, fmaxf(fabsf(o[0]), fmaxf(fabsf(o[1]), fabsf(o[2]))));
    if (pr > best) best = pr, bp[0] = o[0], bp[1] = o[1], bp[2] = o[2];
  }
  for (int off = 32; off > 0; off >>= 1) {
    const float ob = __shfl_xor(best, off);
    const float o0 = __shfl_xor(bp[0], off), o1 = __shfl_xor(bp[1], off), o2 = __shfl_xor(bp[2], off);
    if (ob > best) best = ob, bp[0] = o0, bp[1] = o1, bp[2] = o2;
  }
  for (int off = 32; off > 0; off >>= 1) mag = fmaxf(mag, __shfl_xor(mag, off));
  pt[0] = bp[0], pt[1] = bp[1], pt[2] = bp[2];
  return sign * best;
}
// Best separation gap of P below Q found (negative when none), and the
// magnitude of the coordinates involved.
__device__ float cert_gap(const CertSet& P, const CertSet& Q, const float* u0, float& mag) {
  float u[3] = {u0[0], u0[1], u0[2]};
  float best = -3.4e38f;
  for (int it = 0; it < 4; ++it) {
    const float nu = sqrtf((u[0] * u[0] + u[1] * u[1]) + u[2] * u[2]);
    if (!(nu > 1e-20f)) break;
    float pp[3], qq[3];
    const float hp = cert_support(P, u, 1.f, pp, mag);
    const float lq = cert_support(Q, u, -1.f, qq, mag);
    best = fmaxf(best, (lq - hp) / nu);
    u[0] = qq[0] - pp[0], u[1] = qq[1] - pp[1], u[2] = qq[2] - pp[2];
  }
  return best;
}
__device__ __forceinline__ void cert_mat(const double* m, const double* origin, float* o) {
  for (int r = 0; r < 3; ++r) {
    o[4 * r] = (float)m[4 * r], o[4 * r + 1] = (float)m[4 * r + 1], o[4 * r + 2] = (float)m[4 * r + 2];
    o[4 * r + 3] = (float)(m[4 * r + 3] - (origin ? origin[r] : 0.0));
  }
}

// One wavefront per pair.  mode 0: the CCD interval (a swept against b in
// world coordinates, a's vertices in b's frame at both ends against b, b's in
// a's frame against a); mode 1: the poses of the discrete phase.  gate: the
// MLP outputs (y0 >= s_gate and y1 >= t_gate), or every pair when null.
__global__ void __launch_bounds__(256) k_certify(int n_pairs, const int32_t* pair_info, const int32_t* body_tree,
                                                 const double* pair_rec, const float* Y, float t_gate, float s_gate,
                                                 const int32_t* hv_off, const float* hv, const float* tree_eps_out,
                                                 int mode, int* cert, unsigned long long* counters) {
  const int lane = lane_id();
  const int wstride = gridDim.x * (blockDim.x >> 6);
  for (int p = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); p < n_pairs; p += wstride) {
    const bool gated = !Y || (Y[2 * p] >= s_gate && Y[2 * p + 1] >= t_gate);
    if (!gated) {
      if (lane == 0) cert[p] = 0;
      continue;
    }
    const int a = pair_info[4 * p], b = pair_info[4 * p + 1];
    const int ta = body_tree[a], tb = body_tree[b];
    const double* pr = pair_rec + (size_t)PR_STRIDE * p;
    const float thr = (float)pr[PR_SEARCH] + tree_eps_out[ta] + tree_eps_out[tb];
    CertSet A, B;
    A.v = hv + 3 * (size_t)hv_off[ta], A.n = hv_off[ta + 1] - hv_off[ta];
    B.v = hv + 3 * (size_t)hv_off[tb], B.n = hv_off[tb + 1] - hv_off[tb];
    bool ok = true;
    float mag = 0.f;
    const double* org = pr + PR_TB0;  // world frame relative to b's start position
    const double o3[3] = {org[3], org[7], org[11]};
    float u[3];
    if (mode == 0) {
      // world, both bodies swept
      cert_mat(pr + PR_TA0, o3, A.m[0]), cert_mat(pr + PR_TA1, o3, A.m[1]), A.poses = 2;
      cert_mat(pr + PR_TB0, o3, B.m[0]), cert_mat(pr + PR_TB1, o3, B.m[1]), B.poses = 2;
      for (int k = 0; k < 3; ++k)
        u[k] = 0.5f * ((B.m[0][4 * k + 3] + B.m[1][4 * k + 3]) - (A.m[0][4 * k + 3] + A.m[1][4 * k + 3]));
      float g = cert_gap(A, B, u, mag);
      ok = g - (1e-6f + 64.f * 5.9604645e-08f * mag) > thr;
      if (ok) {  // a's vertices in b's frame at both ends against b
        cert_mat(pr + PR_MAB0, nullptr, A.m[0]), cert_mat(pr + PR_MAB1, nullptr, A.m[1]), A.poses = 2;
        const double I[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
        cert_mat(I, nullptr, B.m[0]), B.poses = 1;
        for (int k = 0; k < 3; ++k) u[k] = -0.5f * (A.m[0][4 * k + 3] + A.m[1][4 * k + 3]);
        mag = 0.f;
        g = cert_gap(A, B, u, mag);
        ok = g - (1e-6f + 64.f * 5.9604645e-08f * mag) > thr;
      }
      if (ok) {  // b's vertices in a's frame at both ends against a
        CertSet A2, B2;
        A2.v = B.v, A2.n = B.n, B2.v = A.v, B2.n = A.n;
        cert_mat(pr + PR_MBA0, nullptr, A2.m[0]), cert_mat(pr + PR_MBA1, nullptr, A2.m[1]), A2.poses = 2;
        const double I[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
        cert_mat(I, nullptr, B2.m[0]), B2.poses = 1;
        for (int k = 0; k < 3; ++k) u[k] = -0.5f * (A2.m[0][4 * k + 3] + A2.m[1][4 * k + 3]);
        mag = 0.f;
        g = cert_gap(A2, B2, u, mag);
        ok = g - (1e-6f + 64.f * 5.9604645e-08f * mag) > thr;
      }
    } else {
      const double* ob = pr + PR_TB1;
      const double o1[3] = {ob[3], ob[7], ob[11]};
      cert_mat(pr + PR_TA1, o1, A.m[0]), A.poses = 1;
      cert_mat(pr + PR_TB1, o1, B.m[0]), B.poses = 1;
      for (int k = 0; k < 3; ++k) u[k] = B.m[0][4 * k + 3] - A.m[0][4 * k + 3];
      const float g = cert_gap(A, B, u, mag);
      ok = g - (1e-6f + 64.f * 5.9604645e-08f * mag) > thr;
    }
    if (lane == 0) {
      cert[p] = ok ? 1 : 0;
      atomicAdd(&counters[0], 1ull);
      if (ok) atomicAdd(&counters[1], 1ull);
    }
  }
}

// The phase's root list without the certified pairs (one block, pair order).
__global__ void __launch_bounds__(1024) k_roots_compact(int n_pairs, const int* cert, NP* soup, unsigned long long* lvl) {
  long long carry = 0;
  for (int t0 = 0; t0 < n_pairs; t0 += blockDim.x) {
    const int p = t0 + threadIdx.x;
    const int keep = (p < n_pairs && !cert[p]) ? 1 : 0;
    int tot;
    const int ex = block_scan(keep, tot);
    if (keep) soup[carry + ex] = NP{(uint32_t)p, 0u, 0u, 0u, 0ull};
    carry += tot;
  }
  if (threadIdx.x == 0) lvl[LV_SOUP] = (unsigned long long)carry;
}
