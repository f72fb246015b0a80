// pw_far.hip -- far field of the tree path: analytic monopole + dipole +
// quadrupole evaluation, the learned residual (MLP on MFMA),
// training-sample generation and the scatter back to the caller's particle
// order.
//
// Units: inside this file velocities and gradients carry a factor 4 pi
// (u' = 4 pi u); the kernels multiply by 1/(4 pi) when they add to the
// rates.
//
// Far field of a node (centre c, moments M, D) at a target x, r = x - c:
//   u'  = M x r / r^3 - w / r^3 + 3 (p x r) / r^5
//   G'  = [M]x / r^3 - 3 (M x r) r^T / r^5
//         + 3 w r^T / r^5 + 3 (D_:m x r + p x e_m)_k / r^5 - 15 (p x r) r^T / r^7
// with w_k = eps_kij D_ij, p = D r, G_km = du_k/dx_m (the singular kernel
// 1/(4 pi r^3) expanded about c to first order in the particle offsets).
// The second-order (quadrupole) term, with Q^(p)_jl = sum a_p delta_j
// delta_l, E_p = Q^(p) r, S_p = r.E_p, P_p = trace Q^(p),
// V_m = eps_mpq E_pq:
//   u'_Q = -3 V / r^5 - 1.5 (P x r) / r^5 + 7.5 (S x r) / r^7,
// and its stretching of a strength a, grad (a.u'_Q):
//   (-3 g - 1.5 a x P) / r^5 + 7.5 (2 sum_j c_j E_j + a x S) / r^7
//   + (15 a.V / r^7 + 7.5 a.(P x r) / r^7 - 52.5 a.(S x r) / r^9) r
// with c = r x a and g_j = sum_m a_m Wq_jm, Wq_jm = eps_mpq Q^(p)_jq.
// The vorticity rate of the target is the transposed contraction
// dalpha_m = sum_k alpha_k G_km, which is the reference's transposed
// stretching q (a_t x a_s) + F [d.(a_t x a_s)] d summed over the node
// (src/interaction/interaction_utils.h:456-501).
//
// The analytic path (no network) evaluates the expansion to quadrupole
// order.  The network path evaluates monopole + dipole, and the network
// stands in for the rest: an MLP 16 -> 128 -> 128 -> 128 -> 12 (GELU, tanh
// form) maps node/target features to the residual (truth - monopole -
// dipole) in scaled units:
//   inputs  e = r l / r^2 (3), M / s0 (3), D / (s0 l) (9), sigma / l (1)
//   outputs u' residual / (s0 l^2 / r^4) (3), G' residual / (s0 l^2 / r^5) (9)
// all expressed in a frame fixed by M and r (make_frame below)
// with l = r_node + sigma the node's length scale.  It is evaluated at the
// centre of each target leaf and carried to the leaf's targets with its own
// gradient output.  Inference runs in BF16
// on the matrix cores with FP32 accumulation, in the transposed form
// H_next^T = W^T H^T: the FP32 result of one 16x16 output tile is exactly
// the B operand of the next layer in the same lanes, so activations never
// leave registers.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "pw_internal.h"
#include "pw_mlp.h"

namespace pw {
namespace {

constexpr double kQd = 0x1.45f306dc9c883p-4;  // 1/(4 pi)
constexpr double kKd = 0x1.0411e71d77959p-4;  // (2 pi)^-1.5

typedef short v4s __attribute__((ext_vector_type(4)));
typedef __bf16 v4bf __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------
// Monopole + dipole of one node at offset r (FP32, 4 pi factored).
// ---------------------------------------------------------------------
__device__ __forceinline__ void md_eval(const float r[3], const float M[3], const float D[9],
                                        float u[3], float G[9]) {
  const float r2 = fmaf(r[2], r[2], fmaf(r[1], r[1], r[0] * r[0]));
  const float ir = __builtin_amdgcn_rsqf(r2);
  const float ir2 = ir * ir;
  const float ir3 = ir * ir2;
  const float ir5 = ir3 * ir2;
  const float ir7 = ir5 * ir2;
  // monopole
  const float mr[3] = {M[1] * r[2] - M[2] * r[1], M[2] * r[0] - M[0] * r[2], M[0] * r[1] - M[1] * r[0]};
  // dipole
  const float w[3] = {D[5] - D[7], D[6] - D[2], D[1] - D[3]};
  const float p[3] = {fmaf(D[2], r[2], fmaf(D[1], r[1], D[0] * r[0])),
                      fmaf(D[5], r[2], fmaf(D[4], r[1], D[3] * r[0])),
                      fmaf(D[8], r[2], fmaf(D[7], r[1], D[6] * r[0]))};
  const float pr[3] = {p[1] * r[2] - p[2] * r[1], p[2] * r[0] - p[0] * r[2], p[0] * r[1] - p[1] * r[0]};
#pragma unroll
  for (int k = 0; k < 3; ++k) u[k] = mr[k] * ir3 - w[k] * ir3 + 3.0f * pr[k] * ir5;
  // [M]x / r^3 + [p]x * 3 / r^5 (the p x e_m term)
  const float a3 = ir3, b5 = 3.0f * ir5;
  G[0] = 0.0f;
  G[1] = -M[2] * a3 - p[2] * b5;
  G[2] = M[1] * a3 + p[1] * b5;
  G[3] = M[2] * a3 + p[2] * b5;
  G[4] = 0.0f;
  G[5] = -M[0] * a3 - p[0] * b5;
  G[6] = -M[1] * a3 - p[1] * b5;
  G[7] = M[0] * a3 + p[0] * b5;
  G[8] = 0.0f;
  // D_:m x r
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    const float c0 = D[m], c1 = D[3 + m], c2 = D[6 + m];
    G[0 + m] += b5 * (c1 * r[2] - c2 * r[1]);
    G[3 + m] += b5 * (c2 * r[0] - c0 * r[2]);
    G[6 + m] += b5 * (c0 * r[1] - c1 * r[0]);
  }
  // rank-one terms: (-3 M x r / r^5 + 3 w / r^5 - 15 p x r / r^7) r^T
  const float c7 = 15.0f * ir7;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float v = b5 * (w[k] - mr[k]) - c7 * pr[k];
    G[3 * k + 0] = fmaf(v, r[0], G[3 * k + 0]);
    G[3 * k + 1] = fmaf(v, r[1], G[3 * k + 1]);
    G[3 * k + 2] = fmaf(v, r[2], G[3 * k + 2]);
  }
}

// Quadrupole slot of the pair (j, l): xx yy zz xy xz yz.
__device__ __forceinline__ constexpr int qslot(int j, int l) {
  return j == l ? j : (j + l == 1 ? 3 : (j + l == 2 ? 4 : 5));
}
// Wq_jm = eps_mpq Q^(p)_jq (row-major, 9) and P_p = trace Q^(p), from Q
__device__ __forceinline__ void quad_derived(const float Q[kQuad], float Wq[9], float P[3]) {
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    Wq[3 * j + 0] = Q[6 * 1 + qslot(j, 2)] - Q[6 * 2 + qslot(j, 1)];
    Wq[3 * j + 1] = Q[6 * 2 + qslot(j, 0)] - Q[6 * 0 + qslot(j, 2)];
    Wq[3 * j + 2] = Q[6 * 0 + qslot(j, 1)] - Q[6 * 1 + qslot(j, 0)];
  }
#pragma unroll
  for (int p = 0; p < 3; ++p) P[p] = Q[6 * p] + Q[6 * p + 1] + Q[6 * p + 2];
}

// Quadrupole of a node at offset r (FP32, 4 pi factored): u += u'_Q when
// with_u, w += grad (a . u'_Q), the stretching of strength a.  ir2 = 1/r^2,
// ir5 = 1/r^5.
__device__ __forceinline__ void quad_accumulate(const float r[3], const float Q[kQuad], const float Wq[9],
                                                const float P[3], const float a[3], float ir2, float ir5,
                                                bool with_u, float u[3], float w[3]) {
  const float ir7 = ir5 * ir2, ir9 = ir7 * ir2;
  float E[3][3], S[3];
#pragma unroll
  for (int p = 0; p < 3; ++p) {
    const float* q = Q + 6 * p;
    E[p][0] = fmaf(q[4], r[2], fmaf(q[3], r[1], q[0] * r[0]));
    E[p][1] = fmaf(q[5], r[2], fmaf(q[1], r[1], q[3] * r[0]));
    E[p][2] = fmaf(q[2], r[2], fmaf(q[5], r[1], q[4] * r[0]));
    S[p] = fmaf(E[p][2], r[2], fmaf(E[p][1], r[1], E[p][0] * r[0]));
  }
  const float V[3] = {E[1][2] - E[2][1], E[2][0] - E[0][2], E[0][1] - E[1][0]};
  const float c5 = -3.0f * ir5, h5 = -1.5f * ir5, c7 = 7.5f * ir7;
  // Z = -1.5 P / r^5 + 7.5 S / r^7, so that u'_Q = -3 V / r^5 + Z x r and
  // the two cross products with a in the stretching are one, a x Z
  const float Z[3] = {fmaf(c7, S[0], h5 * P[0]), fmaf(c7, S[1], h5 * P[1]), fmaf(c7, S[2], h5 * P[2])};
  if (with_u) {
    u[0] = fmaf(c5, V[0], fmaf(Z[1], r[2], fmaf(-Z[2], r[1], u[0])));
    u[1] = fmaf(c5, V[1], fmaf(Z[2], r[0], fmaf(-Z[0], r[2], u[1])));
    u[2] = fmaf(c5, V[2], fmaf(Z[0], r[1], fmaf(-Z[1], r[0], u[2])));
  }
  const float cr[3] = {r[1] * a[2] - r[2] * a[1], r[2] * a[0] - r[0] * a[2], r[0] * a[1] - r[1] * a[0]};  // r x a
  const float f1 = fmaf(a[2], V[2], fmaf(a[1], V[1], a[0] * V[0]));      // a . V
  const float f2 = fmaf(P[2], cr[2], fmaf(P[1], cr[1], P[0] * cr[0]));  // a . (P x r)
  const float f3 = fmaf(S[2], cr[2], fmaf(S[1], cr[1], S[0] * cr[0]));  // a . (S x r)
  const float aZ[3] = {a[1] * Z[2] - a[2] * Z[1], a[2] * Z[0] - a[0] * Z[2], a[0] * Z[1] - a[1] * Z[0]};
  const float rad = fmaf(15.0f * ir7, f1, fmaf(7.5f * ir7, f2, -52.5f * ir9 * f3));
  const float c72 = 2.0f * c7;
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    const float g = fmaf(a[2], Wq[3 * m + 2], fmaf(a[1], Wq[3 * m + 1], a[0] * Wq[3 * m]));
    const float cE = fmaf(cr[2], E[2][m], fmaf(cr[1], E[1][m], cr[0] * E[0][m]));
    w[m] += fmaf(rad, r[m], fmaf(c5, g, fmaf(c72, cE, aZ[m])));
  }
}

// Node data staged in LDS for the far kernels.
// ---------------------------------------------------------------------
// Analytic far field: one thread per target of a leaf, far nodes staged
// through LDS in chunks, positions FP32 relative to the leaf's centre (the
// rounding is 2^-24 of a far node's distance, 1e-7 of the offset).  The
// stretching is contracted with the target's strength node by node,
// dalpha_m = sum_k a_k G[k][m], which for the monopole + dipole field is
//   a x c + (3/r^5) D^T (r x a) + (a.v) r,  c = M/r^3 + 3 (D r)/r^5,
//   v = (3/r^5)(W - M x r) - (15/r^7)((D r) x r),
// W = (D21 - D12, D02 - D20, D10 - D01): the rows of G are never formed.
// ---------------------------------------------------------------------
constexpr int kNodeChunk = 64;

struct FarNodeMD {
  float rc[3];       // centre relative to the target leaf's centre
  float M[3], W[3];  // monopole, dipole's antisymmetric part
  float D[9];
  float Q[kQuad], Wq[9], P[3];  // quadrupole and the parts derived from it
};

template <bool Quad>
__device__ __forceinline__ void stage_md(const NodeRec& q, const NodeQuad* qq, double tcx, double tcy, double tcz,
                                         FarNodeMD& f) {
  f.rc[0] = static_cast<float>(q.cx - tcx);
  f.rc[1] = static_cast<float>(q.cy - tcy);
  f.rc[2] = static_cast<float>(q.cz - tcz);
#pragma unroll
  for (int k = 0; k < 3; ++k) f.M[k] = q.M[k];
#pragma unroll
  for (int k = 0; k < 9; ++k) f.D[k] = q.D[k];
  f.W[0] = q.D[5] - q.D[7];
  f.W[1] = q.D[6] - q.D[2];
  f.W[2] = q.D[1] - q.D[3];
  if (Quad) {
#pragma unroll
    for (int k = 0; k < kQuad; ++k) f.Q[k] = qq->Q[k];
    quad_derived(f.Q, f.Wq, f.P);
  }
}

// u += field of node n at offset r, w += its stretching of strength a
// (4 pi factored; the same terms as md_eval, then with Quad the quadrupole)
template <bool Quad>
__device__ __forceinline__ void md_accumulate(const float r[3], const FarNodeMD& n, const float a[3], float u[3],
                                              float w[3]) {
  const float r2 = fmaf(r[2], r[2], fmaf(r[1], r[1], r[0] * r[0]));
  const float ir = __builtin_amdgcn_rsqf(r2);
  const float ir2 = ir * ir;
  const float a3 = ir * ir2;
  const float b5 = 3.0f * a3 * ir2;
  const float c7 = 5.0f * b5 * ir2;
  const float* M = n.M;
  const float* D = n.D;
  const float p[3] = {fmaf(D[2], r[2], fmaf(D[1], r[1], D[0] * r[0])),
                      fmaf(D[5], r[2], fmaf(D[4], r[1], D[3] * r[0])),
                      fmaf(D[8], r[2], fmaf(D[7], r[1], D[6] * r[0]))};
  // c = M/r^3 + 3 p/r^5: u' = c x r - W/r^3 (the monopole's and the dipole's
  // cross products with r in one)
  const float c[3] = {fmaf(M[0], a3, p[0] * b5), fmaf(M[1], a3, p[1] * b5), fmaf(M[2], a3, p[2] * b5)};
  u[0] = fmaf(-a3, n.W[0], fmaf(c[1], r[2], fmaf(-c[2], r[1], u[0])));
  u[1] = fmaf(-a3, n.W[1], fmaf(c[2], r[0], fmaf(-c[0], r[2], u[1])));
  u[2] = fmaf(-a3, n.W[2], fmaf(c[0], r[1], fmaf(-c[1], r[0], u[2])));
  const float sx = r[1] * a[2] - r[2] * a[1], sy = r[2] * a[0] - r[0] * a[2], sz = r[0] * a[1] - r[1] * a[0];
  // a.v with v = (3/r^5)(W - M x r) - (15/r^7)(p x r): a.(X x r) = X.(r x a)
  const float Y[3] = {fmaf(M[0], b5, p[0] * c7), fmaf(M[1], b5, p[1] * c7), fmaf(M[2], b5, p[2] * c7)};
  const float aW = fmaf(a[2], n.W[2], fmaf(a[1], n.W[1], a[0] * n.W[0]));
  const float av = fmaf(b5, aW, -fmaf(Y[2], sz, fmaf(Y[1], sy, Y[0] * sx)));
  w[0] += fmaf(av, r[0], fmaf(a[1], c[2], fmaf(-a[2], c[1], b5 * fmaf(D[6], sz, fmaf(D[3], sy, D[0] * sx)))));
  w[1] += fmaf(av, r[1], fmaf(a[2], c[0], fmaf(-a[0], c[2], b5 * fmaf(D[7], sz, fmaf(D[4], sy, D[1] * sx)))));
  w[2] += fmaf(av, r[2], fmaf(a[0], c[1], fmaf(-a[1], c[0], b5 * fmaf(D[8], sz, fmaf(D[5], sy, D[2] * sx)))));
  if (Quad) quad_accumulate(r, n.Q, n.Wq, n.P, a, ir2, a3 * ir2, true, u, w);
}

struct FarArgs {
  const Rec64* sorted;
  const NodeRec* nodes;
  const NodeQuad* quad;  // the analytic path's quadrupoles (nullptr on the network path)
  const int* list;
  const int* cnt;
  int cap;
  Rate64* rates;
};

template <bool Quad>
__global__ __launch_bounds__(256) void k_far_md(FarArgs a) {
  __shared__ FarNodeMD sn[kNodeChunk];
  const int T = blockIdx.x;
  const int t0 = a.nodes[T].first, t1 = a.nodes[T].last;
  const double tcx = a.nodes[T].cx, tcy = a.nodes[T].cy, tcz = a.nodes[T].cz;
  // a wave with at most 32 targets gives lanes 32-63 the targets of lanes
  // 0-31 and each half every other node; the halves' sums are added at the end
  const int w0 = t0 + static_cast<int>(threadIdx.x & ~63);  // this wave's first target
  const bool halves = t1 - w0 > 0 && t1 - w0 <= 32;          // wave-uniform
  const int h = halves ? static_cast<int>(threadIdx.x >> 5) & 1 : 0;
  const int t = w0 + static_cast<int>(halves ? threadIdx.x & 31 : threadIdx.x & 63);
  const bool active = t < t1;
  const Rec64 me = a.sorted[active ? t : t0];
  const float rt[3] = {static_cast<float>(me.x - tcx), static_cast<float>(me.y - tcy),
                       static_cast<float>(me.z - tcz)};
  const float at[3] = {static_cast<float>(me.ax), static_cast<float>(me.ay), static_cast<float>(me.az)};
  float u[3] = {0, 0, 0}, w[3] = {0, 0, 0};
  const int nf = a.cnt[T];
  const int* __restrict__ L = a.list + static_cast<size_t>(T) * a.cap;
  for (int base = 0; base < nf; base += kNodeChunk) {
    const int c = min(kNodeChunk, nf - base);
    __syncthreads();
    if (static_cast<int>(threadIdx.x) < c) {
      const int id = L[base + threadIdx.x];
      stage_md<Quad>(a.nodes[id], Quad ? a.quad + id : nullptr, tcx, tcy, tcz, sn[threadIdx.x]);
    }
    __syncthreads();
    if (!active) continue;
    if (halves) {
      for (int k = h; k < c; k += 2) {
        const float r[3] = {rt[0] - sn[k].rc[0], rt[1] - sn[k].rc[1], rt[2] - sn[k].rc[2]};
        md_accumulate<Quad>(r, sn[k], at, u, w);
      }
    } else {
      for (int k = 0; k < c; ++k) {
        const float r[3] = {rt[0] - sn[k].rc[0], rt[1] - sn[k].rc[1], rt[2] - sn[k].rc[2]};
        md_accumulate<Quad>(r, sn[k], at, u, w);
      }
    }
  }
  if (halves) {  // every lane of the wave runs the shuffles
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      u[k] += __shfl_down(u[k], 32);
      w[k] += __shfl_down(w[k], 32);
    }
  }
  if (!active || h != 0) return;
  Rate64 r = a.rates[t];
  r.ux += kQd * u[0];
  r.uy += kQd * u[1];
  r.uz += kQd * u[2];
  r.wx += kQd * w[0];
  r.wy += kQd * w[1];
  r.wz += kQd * w[2];
  a.rates[t] = r;
}

// ---------------------------------------------------------------------
// MLP inference on the matrix cores.
//
// Layer 1 (K = 16): v_mfma_f32_16x16x16_bf16.  Layers 2-4 (K = 128):
// v_mfma_f32_16x16x32_bf16 over four K blocks; K block b pairs the output
// tiles 2b and 2b+1 of the previous layer (slots 0-3 of a lane from tile
// 2b, slots 4-7 from tile 2b+1), and the packed weights use the same
// order of k.  Lane l (c = l % 16, g = l / 16) of an output tile holds rows
// 4g..4g+3 of sample column c.
// ---------------------------------------------------------------------
__device__ __forceinline__ float gelu(float x) {
  // x * sigmoid(2 z), z = sqrt(2/pi) (x + 0.044715 x^3): the tanh form, as
  // x / (1 + 2^(x (c1 + c2 x^2))) with log2(e) folded into c1, c2, on the
  // hardware exp2 and reciprocal (1 ulp each; an IEEE division here cost
  // twice the rest of the network pass)
  constexpr float c1 = -2.0f * 0.7978845608f * 1.4426950409f;
  constexpr float c2 = c1 * 0.044715f;
  const float e = __builtin_amdgcn_exp2f(x * fmaf(c2, x * x, c1));
  return x * __builtin_amdgcn_rcpf(1.0f + e);
}

struct MlpLds {
  v4bf w1[8][64];          // layer 1: 8 output tiles
  v8bf w2[8][4][64];       // layer 2: output tile, K block
  v8bf w3[8][4][64];
  v8bf w4[4][64];          // layer 4: one output tile (12 used)
  float b[kMlpBiasLen];    // b1 | b2 | b3 | b4 (padded to 16)
};

__device__ void load_mlp(const MlpWeights& W, MlpLds& s) {
  const uint4* src = reinterpret_cast<const uint4*>(W.frag);
  uint4* dst = reinterpret_cast<uint4*>(&s);
  constexpr int n16 = kMlpFragBytes / 16;
  for (int i = threadIdx.x; i < n16; i += blockDim.x) dst[i] = src[i];
  for (int i = threadIdx.x; i < kMlpBiasLen; i += blockDim.x) s.b[i] = W.bias[i];
}

// One pass of 32 samples (two 16-column blocks) through the network.
// x[nb] = this lane's 4 input features of column block nb; out[nb] = rows
// 4g..4g+3 of the 12 (16) outputs.
__device__ __forceinline__ void mlp_pass(const MlpLds& s, const v4bf x[2], v4f out[2]) {
  // The weights are the same for every pass; an opaque lane index keeps the
  // compiler from hoisting all 73 KB of fragment reads out of a caller's
  // loop into (spilled) registers: they are read from LDS at their use.
  int lane = threadIdx.x & 63;
  asm volatile("" : "+v"(lane));
  const int g = lane >> 4;
  v4bf h[8][2];
  // layer 1
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const v4s a = __builtin_bit_cast(v4s, s.w1[m][lane]);
#pragma unroll
    for (int nb = 0; nb < 2; ++nb) {
      v4f acc;
#pragma unroll
      for (int i = 0; i < 4; ++i) acc[i] = s.b[16 * m + 4 * g + i];
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, __builtin_bit_cast(v4s, x[nb]), acc, 0, 0, 0);
#pragma unroll
      for (int i = 0; i < 4; ++i) h[m][nb][i] = static_cast<__bf16>(gelu(acc[i]));
    }
  }
  // layers 2 and 3
#pragma unroll
  for (int layer = 0; layer < 2; ++layer) {
    const v8bf(*w)[4][64] = layer == 0 ? s.w2 : s.w3;
    const float* bias = s.b + 128 * (layer + 1);
    v4f acc[8][2];
#pragma unroll
    for (int m = 0; m < 8; ++m)
#pragma unroll
      for (int nb = 0; nb < 2; ++nb)
#pragma unroll
        for (int i = 0; i < 4; ++i) acc[m][nb][i] = bias[16 * m + 4 * g + i];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      v8bf bx[2];
#pragma unroll
      for (int nb = 0; nb < 2; ++nb)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          bx[nb][i] = h[2 * b][nb][i];
          bx[nb][4 + i] = h[2 * b + 1][nb][i];
        }
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        const v8bf a = w[m][b][lane];
#pragma unroll
        for (int nb = 0; nb < 2; ++nb)
          acc[m][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bx[nb], acc[m][nb], 0, 0, 0);
      }
    }
#pragma unroll
    for (int m = 0; m < 8; ++m)
#pragma unroll
      for (int nb = 0; nb < 2; ++nb)
#pragma unroll
        for (int i = 0; i < 4; ++i) h[m][nb][i] = static_cast<__bf16>(gelu(acc[m][nb][i]));
  }
  // layer 4
#pragma unroll
  for (int nb = 0; nb < 2; ++nb)
#pragma unroll
    for (int i = 0; i < 4; ++i) out[nb][i] = s.b[384 + 4 * g + i];
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    const v8bf a = s.w4[b][lane];
#pragma unroll
    for (int nb = 0; nb < 2; ++nb) {
      v8bf bx;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        bx[i] = h[2 * b][nb][i];
        bx[4 + i] = h[2 * b + 1][nb][i];
      }
      out[nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bx, out[nb], 0, 0, 0);
    }
  }
}

// The network works in a frame attached to the node and the offset:
//   e1 = M / |M| (r / |r| when |M| < 1e-3 s0),
//   e2 = the part of r normal to e1, normalised (any normal if r || e1),
//   e3 = e1 x e2;
// R has rows e1, e2, e3.  Inputs are rotated into it (vectors R v, the
// dipole R D R^T) and outputs rotated back (u = R^T u', G = R^T G' R).  The
// residual of a node depends on its shape relative to M and to the offset,
// which this frame fixes, so the network does not have to learn rotations.
__device__ __forceinline__ void make_frame(const float r[3], const float M[3], float s0, float R[9]) {
  const float m2 = fmaf(M[2], M[2], fmaf(M[1], M[1], M[0] * M[0]));
  const float r2 = fmaf(r[2], r[2], fmaf(r[1], r[1], r[0] * r[0]));
  float e1[3];
  if (m2 > 1e-6f * s0 * s0) {
    const float im = 1.0f / sqrtf(m2);
    e1[0] = M[0] * im; e1[1] = M[1] * im; e1[2] = M[2] * im;
  } else {
    const float ir = 1.0f / sqrtf(r2);
    e1[0] = r[0] * ir; e1[1] = r[1] * ir; e1[2] = r[2] * ir;
  }
  const float re = fmaf(r[2], e1[2], fmaf(r[1], e1[1], r[0] * e1[0]));
  float v[3] = {r[0] - re * e1[0], r[1] - re * e1[1], r[2] - re * e1[2]};
  float v2 = fmaf(v[2], v[2], fmaf(v[1], v[1], v[0] * v[0]));
  if (!(v2 > 1e-12f * r2)) {
    const float ax = fabsf(e1[0]) < 0.9f ? 1.0f : 0.0f, ay = 1.0f - ax;
    const float ae = ax * e1[0] + ay * e1[1];
    v[0] = ax - ae * e1[0]; v[1] = ay - ae * e1[1]; v[2] = -ae * e1[2];
    v2 = fmaf(v[2], v[2], fmaf(v[1], v[1], v[0] * v[0]));
  }
  const float iv = 1.0f / sqrtf(v2);
  const float e2[3] = {v[0] * iv, v[1] * iv, v[2] * iv};
  R[0] = e1[0]; R[1] = e1[1]; R[2] = e1[2];
  R[3] = e2[0]; R[4] = e2[1]; R[5] = e2[2];
  R[6] = e1[1] * e2[2] - e1[2] * e2[1];
  R[7] = e1[2] * e2[0] - e1[0] * e2[2];
  R[8] = e1[0] * e2[1] - e1[1] * e2[0];
}

// All 16 inputs of one (point, node) sample at offset r = x - c, the frame,
// and the output scales: u' scale s0 l^2 / r^4, G' scale s0 l^2 / r^5 (those
// of the quadrupole, the first term monopole + dipole leave out).
__device__ __forceinline__ void sample_inputs(const NodeRec& q, const float r[3], float feat[16], float R[9],
                                              float& su, float& sG) {
  const float ell = q.r + q.sg;
  const float d2 = fmaf(r[2], r[2], fmaf(r[1], r[1], r[0] * r[0]));
  const float id2 = 1.0f / d2;
  const float id = __builtin_amdgcn_sqrtf(id2);
  su = q.s0 * ell * ell * id2 * id2;
  sG = su * id;
  make_frame(r, q.M, q.s0, R);
  const float el = ell * id2;
  const float is0 = q.s0 > 0.0f ? 1.0f / q.s0 : 0.0f;
  const float isl = is0 / ell;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    feat[i] = (R[3 * i] * r[0] + R[3 * i + 1] * r[1] + R[3 * i + 2] * r[2]) * el;
    feat[3 + i] = (R[3 * i] * q.M[0] + R[3 * i + 1] * q.M[1] + R[3 * i + 2] * q.M[2]) * is0;
  }
  // R D R^T
  float RD[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) RD[3 * i + j] = R[3 * i] * q.D[j] + R[3 * i + 1] * q.D[3 + j] + R[3 * i + 2] * q.D[6 + j];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      feat[6 + 3 * i + j] = (RD[3 * i] * R[3 * j] + RD[3 * i + 1] * R[3 * j + 1] + RD[3 * i + 2] * R[3 * j + 2]) * isl;
  feat[15] = q.sg / ell;
}

// u = R^T u', G = R^T G' R, scaled, added to acc (12)
__device__ __forceinline__ void unrotate_add(const float* y, const float R[9], float su, float sG, float acc[12]) {
#pragma unroll
  for (int i = 0; i < 3; ++i) acc[i] += su * (R[i] * y[0] + R[3 + i] * y[1] + R[6 + i] * y[2]);
  float GR[9];  // G' R
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int m = 0; m < 3; ++m) GR[3 * a + m] = y[3 + 3 * a] * R[m] + y[4 + 3 * a] * R[3 + m] + y[5 + 3 * a] * R[6 + m];
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int m = 0; m < 3; ++m)
      acc[3 + 3 * k + m] += sG * (R[k] * GR[m] + R[3 + k] * GR[3 + m] + R[6 + k] * GR[6 + m]);
}

// Learned residual over the far lists, evaluated once per (target leaf,
// far node) at the leaf centre c_T and carried to the leaf's targets with
// the gradient the network also predicts:
//   u_res(x_t) = U + G (x_t - c_T),   G_res(x_t) = G,
// U, G = sum over the far nodes of the outputs, rotated back from each
// node's frame and scaled.  The residual falls off as d^-4, so over a leaf
// (r_T < theta d / 2) the first-order transfer keeps it to O((r_T/d)^2)
// relative to itself, while the network runs 256 times less often than per
// target.  8 waves per leaf (one block per CU fits the LDS, so the waves
// of one block hide each other's matrix-core latency); wave w takes the
// 32-node passes w, w+8, ...;
// node k of a pass is sample column k.  The 12 outputs of a sample come out
// spread over four lanes; they are exchanged through LDS so that one lane
// per sample rotates them back.
constexpr int kMlpThreads = 512;
constexpr int kMlpWaves = kMlpThreads / 64;
constexpr int kXch = 24;  // floats exchanged per sample: 12 outputs, frame, 2 scales

__global__ __launch_bounds__(kMlpThreads) void k_far_mlp(FarArgs a, MlpWeights W) {
  extern __shared__ __align__(16) unsigned char smem[];
  MlpLds& s = *reinterpret_cast<MlpLds*>(smem);
  // per wave: 32 samples x {12 outputs, 9 frame, 2 scales}; then the sums
  float* xch = reinterpret_cast<float*>(smem + sizeof(MlpLds));
  float* red = xch + kMlpWaves * 32 * kXch;
  load_mlp(W, s);
  __syncthreads();
  const int T = blockIdx.x;
  const NodeRec leaf = a.nodes[T];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int g = lane >> 4, c = lane & 15;
  float* mine = xch + wave * 32 * kXch;
  const int nf = a.cnt[T];
  const int* __restrict__ L = a.list + static_cast<size_t>(T) * a.cap;
  float acc[12];
#pragma unroll
  for (int i = 0; i < 12; ++i) acc[i] = 0.0f;
  for (int base = 32 * wave; base < nf; base += 32 * kMlpWaves) {
    v4bf x[2];
#pragma unroll
    for (int nb = 0; nb < 2; ++nb) {
      const int j = 16 * nb + c;  // sample column
      const int k = base + j;
      float feat[16], R[9], su = 0.0f, sG = 0.0f;
      if (k < nf) {
        const NodeRec& q = a.nodes[L[k]];
        const float r[3] = {static_cast<float>(leaf.cx - q.cx), static_cast<float>(leaf.cy - q.cy),
                            static_cast<float>(leaf.cz - q.cz)};
        sample_inputs(q, r, feat, R, su, sG);
      } else {
#pragma unroll
        for (int i = 0; i < 16; ++i) feat[i] = 0.0f;
#pragma unroll
        for (int i = 0; i < 9; ++i) R[i] = 0.0f;  // padding column contributes nothing
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) x[nb][i] = static_cast<__bf16>(feat[4 * g + i]);
      if (g == 0) {
#pragma unroll
        for (int i = 0; i < 9; ++i) mine[j * kXch + 12 + i] = R[i];
        mine[j * kXch + 21] = su;
        mine[j * kXch + 22] = sG;
      }
    }
    v4f o[2];
    mlp_pass(s, x, o);
#pragma unroll
    for (int nb = 0; nb < 2; ++nb)
#pragma unroll
      for (int i = 0; i < 4; ++i)
        if (4 * g + i < 12) mine[(16 * nb + c) * kXch + 4 * g + i] = o[nb][i];
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    if (lane < 32) {
      const float* e = mine + lane * kXch;
      unrotate_add(e, e + 12, e[21], e[22], acc);
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  }
  // sum over the 32 sample lanes, then over the waves (in order)
#pragma unroll
  for (int off = 1; off < 32; off <<= 1)
#pragma unroll
    for (int i = 0; i < 12; ++i) acc[i] += __shfl_xor(acc[i], off, 64);
  if (lane == 0)
#pragma unroll
    for (int i = 0; i < 12; ++i) red[wave * 16 + i] = acc[i];
  __syncthreads();
  if (threadIdx.x < 12) {
    float tot = 0.0f;
    for (int w = 0; w < kMlpWaves; ++w) tot += red[16 * w + threadIdx.x];
    red[16 * kMlpWaves + threadIdx.x] = tot;
  }
  __syncthreads();
  const int t = leaf.first + static_cast<int>(threadIdx.x);
  if (t >= leaf.last) return;
  if (threadIdx.x >= kLeafSize) return;
  const float* S = red + 16 * kMlpWaves;  // U (0..2), G (3..11), row-major G[k][m] = S[3 + 3k + m]
  const Rec64 me = a.sorted[t];
  const double dx = me.x - leaf.cx, dy = me.y - leaf.cy, dz = me.z - leaf.cz;
  Rate64 r = a.rates[t];
  r.ux += kQd * (S[0] + S[3] * dx + S[4] * dy + S[5] * dz);
  r.uy += kQd * (S[1] + S[6] * dx + S[7] * dy + S[8] * dz);
  r.uz += kQd * (S[2] + S[9] * dx + S[10] * dy + S[11] * dz);
  r.wx += kQd * (me.ax * S[3] + me.ay * S[6] + me.az * S[9]);
  r.wy += kQd * (me.ax * S[4] + me.ay * S[7] + me.az * S[10]);
  r.wz += kQd * (me.ax * S[5] + me.ay * S[8] + me.az * S[11]);
  a.rates[t] = r;
}

// Plain inference over given inputs (n x 16 FP32, scaled units) -> n x 12:
// the same network pass as the far kernel.
__global__ __launch_bounds__(256) void k_mlp_forward(const float* __restrict__ X, int n,
                                                     float* __restrict__ Y, MlpWeights W) {
  extern __shared__ __align__(16) unsigned char smem[];
  MlpLds& s = *reinterpret_cast<MlpLds*>(smem);
  load_mlp(W, s);
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int g = lane >> 4, c = lane & 15;
  const int base = (blockIdx.x * 4 + wave) * 32;
  if (base >= n) return;
  v4bf x[2];
#pragma unroll
  for (int nb = 0; nb < 2; ++nb) {
    const int j = base + 16 * nb + c;
#pragma unroll
    for (int i = 0; i < 4; ++i) x[nb][i] = static_cast<__bf16>(j < n ? X[16 * j + 4 * g + i] : 0.0f);
  }
  v4f o[2];
  mlp_pass(s, x, o);
#pragma unroll
  for (int nb = 0; nb < 2; ++nb) {
    const int j = base + 16 * nb + c;
    if (j >= n) continue;
#pragma unroll
    for (int i = 0; i < 4; ++i)
      if (4 * g + i < 12) Y[12 * j + 4 * g + i] = o[nb][i];
  }
}

// ---------------------------------------------------------------------
// Training samples: sample k = (target t of leaf T, the node at position
// slot of T's far list).  Exact node sum at the target in FP64 with the
// reference's Gaussian kernel (gaussian.h:47-60), minus the FP32 monopole +
// dipole, in the scaled units of the network (its target); the error of
// the expansion to quadrupole order goes to the side data (the analytic
// path's calibration).
// ---------------------------------------------------------------------
__device__ __forceinline__ void exact_pair(double dx, double dy, double dz, double sig2, const Rec64& s,
                                           double u[3], double G[9]) {
  const double r2 = dx * dx + dy * dy + dz * dz;
  const double rho = sqrt(r2);
  const double sig = sqrt(sig2);
  const double rb = rho / sig;
  const double Z = kKd * exp(-0.5 * rb * rb) / (sig2 * sig);
  const double phi = kQd * erf(rb * 0.70710678118654752440) / (sig2 * sig);
  const double q = (phi / rb - Z) / (rb * rb);
  const double F = (Z - 3.0 * q) / r2;
  const double ax = s.ax, ay = s.ay, az = s.az;
  const double cx = ay * dz - az * dy, cy = az * dx - ax * dz, cz = ax * dy - ay * dx;  // a x d
  u[0] += q * cx; u[1] += q * cy; u[2] += q * cz;
  // G_km = q eps_kim a_i + F d_m (a x d)_k
  const double d[3] = {dx, dy, dz}, ad[3] = {cx, cy, cz};
  G[1] += -q * az; G[2] += q * ay;
  G[3] += q * az;  G[5] += -q * ax;
  G[6] += -q * ay; G[7] += q * ax;
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int m = 0; m < 3; ++m) G[3 * k + m] += F * d[m] * ad[k];
}

struct SampleArgs {
  const Rec64* sorted;
  const NodeRec* nodes;
  const NodeQuad* quad;
  const int* far_list;
  int cap;
  const int* spec;  // per sample: leaf T, list slot, target (sorted index)
  int nsamples;
  float* X;         // nsamples x 16
  float* Y;         // nsamples x 12
  double* aux;      // nsamples x kSampleAux
};

__global__ __launch_bounds__(256) void k_ff_samples(SampleArgs a) {
  __shared__ double red[4 * 12];
  const int k = blockIdx.x;
  const int T = a.spec[3 * k], slot = a.spec[3 * k + 1], t = a.spec[3 * k + 2];
  const int id = a.far_list[static_cast<size_t>(T) * a.cap + slot];
  const NodeRec q = a.nodes[id];
  const Rec64 me = a.sorted[t];
  double v[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int s = q.first + static_cast<int>(threadIdx.x); s < q.last; s += 256) {
    const Rec64 src = a.sorted[s];
    exact_pair(me.x - src.x, me.y - src.y, me.z - src.z, 0.5 * (me.s2 + src.s2), src, v, v + 3);
  }
  for (int off = 32; off > 0; off >>= 1)
#pragma unroll
    for (int m = 0; m < 12; ++m) v[m] += __shfl_xor(v[m], off, 64);
  if ((threadIdx.x & 63) == 0)
    for (int m = 0; m < 12; ++m) red[(threadIdx.x >> 6) * 12 + m] = v[m];
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int m = 0; m < 12; ++m) v[m] = ((red[m] + red[12 + m]) + red[24 + m]) + red[36 + m];
  const float r[3] = {static_cast<float>(me.x - q.cx), static_cast<float>(me.y - q.cy),
                      static_cast<float>(me.z - q.cz)};
  float uu[3], gg[9];
  md_eval(r, q.M, q.D, uu, gg);
  float uq[3] = {uu[0], uu[1], uu[2]};  // + the quadrupole's velocity
  {
    const NodeQuad qq = a.quad[id];
    float Wq[9], P[3], wq[3] = {0.0f, 0.0f, 0.0f};
    quad_derived(qq.Q, Wq, P);
    const float r2 = fmaf(r[2], r[2], fmaf(r[1], r[1], r[0] * r[0]));
    const float ir2 = 1.0f / r2;
    const float ir5 = ir2 * ir2 * sqrtf(ir2);
    const float e0[3] = {1.0f, 0.0f, 0.0f};
    quad_accumulate(r, qq.Q, Wq, P, e0, ir2, ir5, true, uq, wq);
  }
  // inputs in FP32 (the network rounds them to BF16 itself) and the frame
  float feat[16], R[9], su, sG;
  sample_inputs(q, r, feat, R, su, sG);
  double res[12], err_md = 0.0, un = 0.0, err_q = 0.0;
  for (int m = 0; m < 12; ++m) {
    const double fourpi_exact = v[m] * 12.566370614359172954;  // 4 pi
    res[m] = fourpi_exact - static_cast<double>(m < 3 ? uu[m] : gg[m - 3]);
    if (m < 3) {
      err_md += res[m] * res[m];
      un += fourpi_exact * fourpi_exact;
      const double eq = fourpi_exact - static_cast<double>(uq[m]);
      err_q += eq * eq;
    }
  }
  // targets in the frame: R res_u / su, R res_G R^T / sG
  float* y = a.Y + 12 * static_cast<size_t>(k);
  for (int i = 0; i < 3; ++i)
    y[i] = static_cast<float>((R[3 * i] * res[0] + R[3 * i + 1] * res[1] + R[3 * i + 2] * res[2]) / su);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double acc = 0.0;
      for (int p = 0; p < 3; ++p)
        for (int q2 = 0; q2 < 3; ++q2) acc += R[3 * i + p] * res[3 + 3 * p + q2] * R[3 * j + q2];
      y[3 + 3 * i + j] = static_cast<float>(acc / sG);
    }
  for (int i = 0; i < 16; ++i) a.X[16 * static_cast<size_t>(k) + i] = feat[i];
  double* o = a.aux + static_cast<size_t>(k) * kSampleAux;
  const double d = sqrt(static_cast<double>(r[0]) * r[0] + static_cast<double>(r[1]) * r[1] +
                        static_cast<double>(r[2]) * r[2]);
  o[0] = su;               // u scale (4 pi units)
  o[1] = sG;               // G scale
  o[2] = d;                // |x - c|
  o[3] = q.s2;             // node s2
  o[4] = sqrt(err_md);     // |u' exact - u' expansion|
  o[5] = sqrt(un);         // |u' exact|
  o[6] = q.r;              // node radius
  o[7] = id;               // node id
  o[8] = t;                // target (Morton-sorted index)
  o[9] = T;                // target leaf
  o[10] = sqrt(err_q);     // |u' exact - u' to quadrupole order|
  o[11] = q.s3;            // node s3 (bound)
}

// Conservation of total vorticity: the pairwise antisymmetry that makes
// sum dalpha vanish in the direct sum (interaction_utils.h:493-495) does not
// hold for one-sided far-field evaluations, so the mean vorticity rate is
// removed: block sums (FP64, fixed order), their total, the subtraction.
// block b sums particles b*256 + t + k*(grid*256) (each lane in k order),
// then its lanes in a fixed order
__global__ __launch_bounds__(256) void k_sum_dalpha(const Rate64* __restrict__ r, int n, double* __restrict__ part) {
  __shared__ double red[4][3];
  double v[3] = {0.0, 0.0, 0.0};
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    v[0] += r[i].wx;
    v[1] += r[i].wy;
    v[2] += r[i].wz;
  }
  for (int off = 32; off > 0; off >>= 1)
    for (int k = 0; k < 3; ++k) v[k] += __shfl_xor(v[k], off, 64);
  if ((threadIdx.x & 63) == 0)
    for (int k = 0; k < 3; ++k) red[threadIdx.x >> 6][k] = v[k];
  __syncthreads();
  if (threadIdx.x < 3)
    part[3 * blockIdx.x + threadIdx.x] =
        (red[0][threadIdx.x] + red[1][threadIdx.x]) + (red[2][threadIdx.x] + red[3][threadIdx.x]);
}

// the mean of the nb block sums: one block, lane t takes blocks t, t+256,
// ..., then the lanes in a fixed order
__global__ __launch_bounds__(256) void k_mean_dalpha(double* __restrict__ part, int nb, int n) {
  __shared__ double red[4][3];
  double v[3] = {0.0, 0.0, 0.0};
  for (int b = threadIdx.x; b < nb; b += 256)
    for (int k = 0; k < 3; ++k) v[k] += part[3 * b + k];
  for (int off = 32; off > 0; off >>= 1)
    for (int k = 0; k < 3; ++k) v[k] += __shfl_xor(v[k], off, 64);
  if ((threadIdx.x & 63) == 0)
    for (int k = 0; k < 3; ++k) red[threadIdx.x >> 6][k] = v[k];
  __syncthreads();
  if (threadIdx.x < 3)
    part[3 * nb + threadIdx.x] = ((red[0][threadIdx.x] + red[1][threadIdx.x]) + (red[2][threadIdx.x] + red[3][threadIdx.x])) /
                                 static_cast<double>(n);
}

__global__ __launch_bounds__(256) void k_sub_dalpha(Rate64* __restrict__ r, int n, const double* __restrict__ mean) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  r[i].wx -= mean[0];
  r[i].wy -= mean[1];
  r[i].wz -= mean[2];
}

// rates[perm[i]] = rates_s[i]
__global__ __launch_bounds__(256) void k_scatter(const Rate64* __restrict__ in, const int* __restrict__ perm,
                                                 int n, Rate64* __restrict__ out) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) out[perm[i]] = in[i];
}

}  // namespace

size_t mlp_far_lds_bytes() { return sizeof(MlpLds) + (kMlpWaves * 32 * kXch + 16 * kMlpWaves + 16) * sizeof(float); }

hipError_t launch_far(const Rec64* sorted, const NodeRec* nodes, const NodeQuad* quad, const int* far_list,
                      const int* far_cnt, int cap_far, int nleaf, const MlpWeights* mlp, Rate64* rates_s,
                      hipStream_t st) {
  hipError_t e = launch_far_md(sorted, nodes, quad, far_list, far_cnt, cap_far, nleaf, rates_s, st);
  if (e != hipSuccess || mlp == nullptr) return e;
  return launch_far_mlp(sorted, nodes, far_list, far_cnt, cap_far, nleaf, *mlp, rates_s, st);
}

hipError_t launch_far_md(const Rec64* sorted, const NodeRec* nodes, const NodeQuad* quad, const int* far_list,
                         const int* far_cnt, int cap_far, int nleaf, Rate64* rates_s, hipStream_t st) {
  if (nleaf <= 0) return hipSuccess;
  FarArgs a{sorted, nodes, quad, far_list, far_cnt, cap_far, rates_s};
  if (quad) hipLaunchKernelGGL(k_far_md<true>, dim3(nleaf), dim3(256), 0, st, a);
  else hipLaunchKernelGGL(k_far_md<false>, dim3(nleaf), dim3(256), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_far_mlp(const Rec64* sorted, const NodeRec* nodes, const int* far_list, const int* far_cnt,
                          int cap_far, int nleaf, const MlpWeights& mlp, Rate64* rates_s, hipStream_t st) {
  if (nleaf <= 0) return hipSuccess;
  FarArgs a{sorted, nodes, nullptr, far_list, far_cnt, cap_far, rates_s};
  hipError_t e;
  const size_t lds = mlp_far_lds_bytes();
  static std::atomic<unsigned long long> lds_raised{0};  // LDS beyond the 64 KB default
  e = ensure_lds_limit(reinterpret_cast<const void*>(k_far_mlp), lds, lds_raised);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_far_mlp, dim3(nleaf), dim3(kMlpThreads), lds, st, a, mlp);
  return hipGetLastError();
}

constexpr int kDalphaBlocks = 1024;  // block sums of the mean vorticity rate at most

hipError_t launch_remove_mean_dalpha(Rate64* rates, int n, double* part, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const int nb = std::min((n + 255) / 256, kDalphaBlocks);  // a function of n only: a fixed summation order
  hipLaunchKernelGGL(k_sum_dalpha, dim3(nb), dim3(256), 0, st, rates, n, part);
  hipLaunchKernelGGL(k_mean_dalpha, dim3(1), dim3(256), 0, st, part, nb, n);
  hipLaunchKernelGGL(k_sub_dalpha, dim3((n + 255) / 256), dim3(256), 0, st, rates, n, part + 3 * nb);
  return hipGetLastError();
}

hipError_t launch_scatter_rates(const Rate64* rates_s, const int* perm, int n, Rate64* rates, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_scatter, dim3((n + 255) / 256), dim3(256), 0, st, rates_s, perm, n, rates);
  return hipGetLastError();
}

hipError_t launch_mlp_forward(const float* X, int n, float* Y, const MlpWeights& W, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const int blocks = (n + 127) / 128;
  static std::atomic<unsigned long long> lds_raised{0};
  const hipError_t e = ensure_lds_limit(reinterpret_cast<const void*>(k_mlp_forward), sizeof(MlpLds), lds_raised);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_mlp_forward, dim3(blocks), dim3(256), sizeof(MlpLds), st, X, n, Y, W);
  return hipGetLastError();
}

hipError_t launch_ff_samples(const Rec64* sorted, const NodeRec* nodes, const NodeQuad* quad, const int* far_list,
                             int cap, const int* spec, int nsamples, float* X, float* Y, double* aux,
                             hipStream_t st) {
  if (nsamples <= 0) return hipSuccess;
  SampleArgs a{sorted, nodes, quad, far_list, cap, spec, nsamples, X, Y, aux};
  hipLaunchKernelGGL(k_ff_samples, dim3(nsamples), dim3(256), 0, st, a);
  return hipGetLastError();
}

}  // namespace pw
