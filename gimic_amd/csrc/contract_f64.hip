// FP64 contraction of basis values with the density family on MFMA, with the
// per-point reductions fused into the epilogue (gfx950).
//
// The reference's contract (jtensor.F90:148-237) does, per point, seven
// length-n GEMVs phi*D, phi*DB_b, db_b*D and 27 dot products.  db_b = S_b phi
// with S = r x R_A (bfeval.f90:168-293), so every GEMV shares the left
// operand phi and a tile of P points is one GEMM
//     Y[P x 7n] = Phi[P x n] * [D | DBx | DBy | DBz | Wx | Wy | Wz]
// followed by reductions over the output column nu against phi_nu,
// grad phi_nu and R_nu (the kNQ partial sums listed in device_common.h).
// The finalize kernel turns the partials into the 3x3 tensor.
//
// With W_d(mu,nu) = D(mu,nu)(R_nu - R_mu)_d in place of R_d D, the GIAO
// reduction needs no difference of large terms, and the diamagnetic term
// is accumulated as sum Y0 phi (r - R_nu) (gauge origin at each function's
// centre), again free of cancellation.
//
// Tiling: WAVES wavefronts (4, or 2 for small launches); a workgroup owns
// 32 WAVES point rows (32 per wavefront: two 16-row MFMA blocks) and walks a
// range of 16-column N-tiles; per N-tile each wavefront keeps 2 x 7
// v_mfma_f64_16x16x4 blocks.  Row tiles of 64 rather than 128 points list
// fewer groups where neighbouring points are far apart, as in the FP64 pass
// over the points next to nuclei, whose 128-row tiles mostly take in two
// nuclei.
// Only what screening leaves non-zero over the row tile (tiles.hip) is
// visited: the live 8-function groups of its listed 32-function blocks,
// two per 16-column N-tile (split evenly over the workgroups that share
// the row tile) and two per 16-wide K-step.
// The K index is remapped inside every K-step so each lane loads 32
// contiguous bytes of its A row and of each B column (lane group g holds
// elements 4(g&1)..+3 of group g/2; sub-step s feeds the s-th of them to
// both operands).
#include "device_common.h"

namespace gimic {

template <int WAVES>
__global__ __launch_bounds__(64 * WAVES) void contract_f64_kernel(ContractArgs a) {
    constexpr int kRows = 32 * WAVES;
    __shared__ double qs[kRows * kNQ];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int g = lane >> 4, c = lane & 15;
    const int set = blockIdx.z, split = blockIdx.y;
    const size_t row0 = (size_t)blockIdx.x * kRows;
    const int lda = a.lda;
    const size_t rows = (size_t)a.rows;
    const size_t matsz = (size_t)a.ncol_pad * lda;
    const size_t gplane = (size_t)a.gplane;
    for (int i = tid; i < kRows * kNQ; i += 64 * WAVES) qs[i] = 0.0;
    __syncthreads();

    const double *A0 = a.a64 + (row0 + wave * 32 + c) * lda;
    const double *A1 = A0 + 16 * (size_t)lda;
    const double *Bset = a.b64 + (size_t)set * kNMat * matsz;
    // the live 8-function groups of the listed blocks (tiles.hip), two per
    // 16-wide K-step: lane group g takes group 2 kk + g / 2, elements
    // 4 (g & 1) .. + 3.  A padding entry (-1) reads A from the zero buffer
    // (B from any valid place: the product is zero).  Two per N-tile too:
    // lane c owns function c & 7 of group 2 e + c / 8 (a padding entry: a
    // column whose products the epilogue drops).
    const int *K8 = a.k8list + (size_t)blockIdx.x * a.nblk * 4;
    const int nks = (a.k8count[blockIdx.x] + 1) >> 1;
    const int e0 = (int)((long long)split * nks / a.nsplit);
    const int e1 = (int)((long long)(split + 1) * nks / a.nsplit);
    const double *zero = (const double *)a.bzero;
    auto kent = [&](int kk) { return K8[2 * kk + (g >> 1)]; };
    auto koff = [&](int v) { return v < 0 ? 0 : (v >> 12) * 8 + (g & 1) * 4; };

    for (int e = e0; e < e1; e++) {
        const int xn = K8[2 * e + (c >> 3)];
        const bool live = xn >= 0;
        const int col = live ? (xn >> 12) * 8 + (c & 7) : 0;
        const double *Bc = Bset + (size_t)col * lda;
        v4d acc[2][kNMat];
#pragma unroll
        for (int m = 0; m < kNMat; m++) acc[0][m] = acc[1][m] = v4d{0.0, 0.0, 0.0, 0.0};

        const int v0 = kent(0), k0 = koff(v0);
        v4d a0 = *(const v4d *)(v0 < 0 ? zero : A0 + k0), a1 = *(const v4d *)(v0 < 0 ? zero : A1 + k0);
        v4d b[kNMat];
#pragma unroll
        for (int m = 0; m < kNMat; m++) b[m] = *(const v4d *)(Bc + m * matsz + k0);
        // group entry of the next K-step, read one step early
        int vnext = kent(nks > 1 ? 1 : 0);
        for (int kk = 0; kk < nks; kk++) {
            // prefetch the next K-step while this one runs on the matrix cores
            const int vn = vnext, kn = koff(vn);
            vnext = kent(kk + 2 < nks ? kk + 2 : nks - 1);
            const v4d na0 = *(const v4d *)(vn < 0 ? zero : A0 + kn);
            const v4d na1 = *(const v4d *)(vn < 0 ? zero : A1 + kn);
            v4d nb[kNMat];
#pragma unroll
            for (int m = 0; m < kNMat; m++) nb[m] = *(const v4d *)(Bc + m * matsz + kn);
#pragma unroll
            for (int s = 0; s < 4; s++) {
#pragma unroll
                for (int m = 0; m < kNMat; m++) {
                    acc[0][m] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[s], b[m][s], acc[0][m], 0, 0, 0);
                    acc[1][m] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[s], b[m][s], acc[1][m], 0, 0, 0);
                }
            }
            a0 = na0;
            a1 = na1;
#pragma unroll
            for (int m = 0; m < kNMat; m++) b[m] = nb[m];
        }

        // epilogue: for the f64 16x16x4 builtin, lane (g, c) holds
        // Y_mat[row g + 4r][col c] of each block (checked by the MFMA
        // layout self-test)
        const double R[3] = {a.rf64[col], a.rf64[a.ncol_pad + col], a.rf64[2 * a.ncol_pad + col]};
#pragma unroll
        for (int rb = 0; rb < 2; rb++) {
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int rl = wave * 32 + rb * 16 + g + 4 * r;
                const size_t p = row0 + rl;
                const bool valid = p < (size_t)a.npts;
                const double rp[3] = {valid ? a.pts[3 * p] : 0.0, valid ? a.pts[3 * p + 1] : 0.0,
                                      valid ? a.pts[3 * p + 2] : 0.0};
                const double ph = live ? a.a64[p * lda + col] : 0.0;
                const double G[3] = {live ? a.g64[p * lda + col] : 0.0,
                                     live ? a.g64[gplane + p * lda + col] : 0.0,
                                     live ? a.g64[2 * gplane + p * lda + col] : 0.0};
                const double y0 = acc[rb][0][r];
                double q[kNQ];
                q[0] = y0 * ph;
#pragma unroll
                for (int d = 0; d < 3; d++) q[1 + d] = q[0] * (rp[d] - R[d]);
#pragma unroll
                for (int bb = 0; bb < 3; bb++)
#pragma unroll
                    for (int m = 0; m < 3; m++) q[4 + m + 3 * bb] = acc[rb][1 + bb][r] * G[m];
#pragma unroll
                for (int d = 0; d < 3; d++)
#pragma unroll
                    for (int m = 0; m < 3; m++) q[13 + m + 3 * d] = G[m] * acc[rb][4 + d][r];
#pragma unroll
                for (int i = 0; i < kNQ; i++) q[i] = row16_sum(q[i]);
#pragma unroll
                for (int i = 0; i < kNQ; i++)
                    if ((i & 15) == c) qs[rl * kNQ + i] += q[i];
            }
        }
    }
    __syncthreads();
    double *out = a.part + ((size_t)set * a.nsplit + split) * kNQ * rows + row0;
    for (int i = tid; i < kRows * kNQ; i += 64 * WAVES) {
        const int q = i / kRows, rl = i % kRows;
        out[(size_t)q * rows + rl] = qs[rl * kNQ + q];
    }
}

hipError_t launch_contract_f64(const ContractArgs &a, int tile_rows, hipStream_t s) {
    if (tile_rows != 64 && tile_rows != 128) return hipErrorInvalidValue;
    if (a.rows % tile_rows) return hipErrorInvalidValue;
    dim3 grid(a.rows / tile_rows, a.nsplit, a.nmatset);
    if (tile_rows == 64)
        hipLaunchKernelGGL(contract_f64_kernel<2>, grid, dim3(128), 0, s, a);
    else
        hipLaunchKernelGGL(contract_f64_kernel<4>, grid, dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace gimic
