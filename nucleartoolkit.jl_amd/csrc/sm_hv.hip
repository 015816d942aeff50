// H·v on gfx950: one workgroup per group of 4 consecutive output proton rows.
//
// Output element (bf, fp, fn) (block, proton SD, neutron SD; global index
// tdims[bf] + fp*nN + fn, the reference's order) is owned by one lane; all
// contributions are gathered, so no atomics are needed and the result does not
// depend on the launch shape.  In the order they are added:
//   pn       : for every input block bi coupled by a proton dM, for every proton
//              hop ip -> fp (operator ac, sign s_p):
//                sum_{neutron hops in -> fn (operator bd, sign s_n)}
//                    s_p s_n V_eff[dM][ac][bd] v[bi: ip, in]
//              The input rows s_p v[ip,:] and the V_eff rows of the hop's
//              operators are staged in LDS; each lane walks its own neutron hop
//              list (sliced ELL: 64 fn per slice, in the block's walk order
//              nperm, coalesced); the sums then move to the natural order
//   diagonal : (e_p[fp] + e_n[fn]) v[fp,fn]
//   nn       : sum_k c_k v[fp, col_k]         (row fp of v staged in LDS)
//   pp       : sum_e c_e v[ip_e, fn]          (coalesced rows of the same block)
// The reference's loop nest for the same sums is operate_H!
// (src/ShellModel/lanczos_methods.jl:153-248).
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <stdexcept>

#include "sm_device.h"

namespace smgpu {

static constexpr size_t HV_LDS_TWO_ROWS = 16 * 1024;  // two-row groups for narrow blocks only

DBuf::~DBuf() {
    if (p && owned) (void)hipFree(p);
}
void DBuf::release() {
    if (p && owned) (void)hipFree(p);
    p = nullptr;
    bytes = cap = 0;
    owned = true;
}
void DBuf::alloc(size_t n) {
    release();
    if (n) SM_HIP(hipMalloc(&p, n));
    bytes = cap = n;
}

DeviceHam::~DeviceHam() {
    if (stream) (void)hipStreamDestroy(stream);
}

template <typename T> struct Vec4;
template <> struct Vec4<float> { using type = float4; };
template <> struct Vec4<double> { using type = double4; };

// One workgroup per group of RG consecutive proton rows of one block (lanes
// along the neutron index, NPT columns per lane), for NV input vectors at once
// (P.v[b] -> P.w[b], b < NV; the reference's bl_operate_H!,
// lanczos_methods.jl:435-548, for NV > 1).  Every neutron-side table entry and
// V_eff slot a lane loads (nn and pn hop lists) is applied to all rows of the
// group and all NV vectors.
// pn: per coupled input block, the proton hops of the group's rows are staged
// four per row per round ("slots"), slot-interleaved in LDS:
//   X[in][b][slot] = s_p v_b[bi: ip_slot, in],   Vs[col][slot] = V_eff[dM][ac_slot][col]
// so that a lane fetches the four hops of a row for its neutron hop (in, col)
// with one 4-wide LDS load per vector and one for V_eff, at compile-time offsets.
// Unused slots are zero.
template <typename T, int NPT, int RG, int PCH, int NV, bool PV, int TPB>
__global__ __launch_bounds__(TPB) void k_hv(HvTables A, HvVecs P) {
    static_assert(PCH % 4 == 0, "slots come in groups of four");
    constexpr int NG = PCH / 4;       // slot groups per row and round
    constexpr int SL = RG * PCH;      // slots per round
    using T4 = typename Vec4<T>::type;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    constexpr int NC = PV ? NV : 1;  // coefficient sets (PV: one per vector)
    constexpr int NW = TPB / 64;     // waves per workgroup
    T* X = reinterpret_cast<T*>(smem_raw);                      // nN_max * NV * SL
    T* Vs = X + (size_t)A.nN_max * SL * NV;                      // NC * cols2_max * SL
    T* xs = Vs + (size_t)A.cols2_max * SL * NC;                  // nN_max * RG * NV  ([col][row][vec])
    const int g = blockIdx.x;
    const int tid = threadIdx.x;
    const int bf = A.grp_block[g];
    const int fp0 = A.grp_fp0[g];
    const int nN = A.nN[bf];
    const int nr = min(RG, A.nP[bf] - fp0);
    const int row0 = A.block_row0[bf] + fp0;          // global row index of fp0
    const int64_t bstart = A.tdims[bf];
    // coefficient set of vector b (the launch's Hamiltonian, or vector b's own for PV)
    auto cep = [&](int b) { return static_cast<const T*>(PV ? P.ep[b] : A.ep); };
    auto cen = [&](int b) { return static_cast<const T*>(PV ? P.en[b] : A.en) + A.en_off[bf]; };
    auto cpp = [&](int b) { return static_cast<const T*>(PV ? P.pp[b] : A.pp_val); };
    auto cnn = [&](int b) { return static_cast<const T*>(PV ? P.nn[b] : A.nn_val); };
    auto cve = [&](int b) { return static_cast<const T*>(PV ? P.veff[b] : A.veff); };
    const int64_t vshift = P.vsel ? (int64_t)*P.vsel * P.vld : 0;
    const T* vin[NV];
#pragma unroll
    for (int b = 0; b < NV; b++) vin[b] = static_cast<const T*>(P.v[b]) + vshift;
    for (int i = tid; i < nN; i += TPB) {
#pragma unroll
        for (int r = 0; r < RG; r++)
#pragma unroll
            for (int b = 0; b < NV; b++)
                xs[(i * RG + r) * NV + b] =
                    r < nr ? vin[b][bstart + (int64_t)(fp0 + r) * nN + i] : T(0);
    }
    __syncthreads();

    T acc[NPT][RG][NV];
    T epr[RG][NC];
#pragma unroll
    for (int r = 0; r < RG; r++)
#pragma unroll
        for (int c = 0; c < NC; c++) epr[r][c] = r < nr ? cep(c)[row0 + r] : T(0);
    const int64_t nsb = A.nn_slice_base[bf];
    // Column k of a lane is fn = tid + TPB k, in slice (tid >> 6) + NW k: the same for
    // the whole wave, so slice tests are wave-uniform.  Lanes past nN inside the
    // last slice work on the clamped column nN - 1 (valid addresses; their results
    // are never stored) and read the zero padding of the ELL tables.
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int nsl = (nN + 63) >> 6;
#pragma unroll
    for (int k = 0; k < NPT; k++)
#pragma unroll
        for (int r = 0; r < RG; r++)
#pragma unroll
            for (int b = 0; b < NV; b++) acc[k][r][b] = T(0);
    // pn
    const int q1 = A.pair_ptr[bf + 1];
    for (int q = A.pair_ptr[bf]; q < q1; q++) {
        const int pt = A.pair_ptab[q];
        const int32_t* prp = A.ptab_rowptr + A.ptab_rowptr_base[pt] + fp0;
        int rb[RG + 1];
#pragma unroll
        for (int r = 0; r <= RG; r++) rb[r] = prp[min(r, nr)];
        int nch = 0;
#pragma unroll
        for (int r = 0; r < RG; r++) nch = max(nch, (rb[r + 1] - rb[r] + PCH - 1) / PCH);
        if (nch == 0) continue;
        const int bi = A.pair_bi[q];
        const int dm = A.pair_dm[q];
        const int nt = A.pair_ntab[q];
        const uint32_t* pent = A.ptab_ent + A.ptab_ent_base[pt];
        const int nNi = A.nN[bi];
        const int64_t ibase = A.tdims[bi];
        const int64_t vbase = A.veff_off[dm];
        const int c2 = A.veff_cols2[dm];
        const int64_t sb = A.ntab_slice_base[nt];
        // column k of a lane is fn = tid + TPB k: slice (tid >> 6) + NW k, the same
        // for the whole wave, so slice lengths (and every test on them) are
        // wave-uniform; lanes past nN inside the last slice read its zero-column
        // padding, waves past the last slice get length 0 and skip all work
        int64_t off[NPT];
        int len[NPT];
        int maxlen = 0;
#pragma unroll
        for (int k = 0; k < NPT; k++) {
            const int sk = wave + NW * k;
            if (sk < nsl) {
                const int64_t s = sb + sk;
                off[k] = A.ntab_slice_off[s] + (tid & 63);
                len[k] = __builtin_amdgcn_readfirstlane(A.ntab_slice_len[s]);
            } else {
                off[k] = A.ntab_slice_off[sb] + (tid & 63);   // a valid row: every slice has one
                len[k] = 0;
            }
            maxlen = max(maxlen, len[k]);
        }
        maxlen = __builtin_amdgcn_readfirstlane(maxlen);
        for (int ch = 0; ch < nch; ch++) {
            __syncthreads();
            // Stage the round's active slot groups (four hops of one row each):
            // X[in][b][slot] = s v_b[bi: ip, in] and Vs[col][slot] = V_eff[op][col].
            // Offsets, not pointers, so every gather is an unconditional global
            // load from a kernel argument (a group's unused slots read a valid
            // element times 0).  Groups past a row's hops are neither staged nor read.
#pragma unroll
            for (int r = 0; r < RG; r++)
#pragma unroll
                for (int g = 0; g < NG; g++) {
                    const int h0 = rb[r] + ch * PCH + g * 4;
                    if (h0 >= rb[r + 1]) continue;                 // uniform
                    int64_t xo[4];
                    int vo[4];
                    T sg[4], vm[4];
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        if (h0 + j < rb[r + 1]) {
                            const uint32_t e = pent[h0 + j];
                            xo[j] = ibase + (int64_t)(e & 0xFFFFFu) * nNi;
                            vo[j] = (int)((e >> 20) & 0x7FFu) * c2;
                            sg[j] = (e >> 31) ? T(-1) : T(1);
                            vm[j] = T(1);
                        } else {
                            xo[j] = ibase; vo[j] = 0; sg[j] = T(0); vm[j] = T(0);
                        }
                    }
                    const int so = r * PCH + g * 4;
                    for (int i = tid; i < nNi; i += TPB) {
#pragma unroll
                        for (int b = 0; b < NV; b++) {
                            const T* vb = vin[b];
                            T4 x4;
                            x4.x = sg[0] * vb[xo[0] + i];
                            x4.y = sg[1] * vb[xo[1] + i];
                            x4.z = sg[2] * vb[xo[2] + i];
                            x4.w = sg[3] * vb[xo[3] + i];
                            *reinterpret_cast<T4*>(X + ((size_t)i * NV + b) * SL + so) = x4;
                        }
                    }
#pragma unroll
                    for (int cc = 0; cc < NC; cc++) {
                        const T* V2 = cve(cc) + vbase;
                        T* Vd = Vs + (size_t)cc * A.cols2_max * SL;
                        for (int c = tid; c < c2; c += TPB) {
                            T4 v4;
                            v4.x = vm[0] * V2[vo[0] + c];
                            v4.y = vm[1] * V2[vo[1] + c];
                            v4.z = vm[2] * V2[vo[2] + c];
                            v4.w = vm[3] * V2[vo[3] + c];
                            *reinterpret_cast<T4*>(Vd + (size_t)c * SL + so) = v4;
                        }
                    }
                }
            __syncthreads();
            if constexpr (NPT <= 4) {
                // Column by column (k outer): a chunk of PF entries of one column is
                // decoded, the next chunk's entry loads are issued, then the chunk's LDS
                // reads are all issued before its FMAs (one LDS wait per chunk).  Per
                // accumulator the order of entries and slots is that of one entry at a
                // time.  PF trades LDS reads in flight against VGPRs (occupancy):
                // measured on Cr48, four entries suit one FP32 vector (0.478 -> 0.447 ms),
                // two everything else (two vectors 0.331 -> 0.315 ms per vector, FP64
                // 0.72 -> 0.70 ms).
                constexpr int PF = (NV == 1 && sizeof(T) == 4) ? 4 : 2;
#pragma unroll
                for (int k = 0; k < NPT; k++) {
                    const int lk = len[k];
                    if (lk == 0) continue;                                // wave-uniform
                    uint32_t raw[PF];
#pragma unroll
                    for (int d = 0; d < PF; d++) raw[d] = A.ntab_ent[off[k] + (int64_t)min(d, lk - 1) * 64];
                    for (int q0 = 0; q0 < lk; q0 += PF) {
                        int in[PF], col[PF];
#pragma unroll
                        for (int d = 0; d < PF; d++) {
                            in[d] = (int)(raw[d] & 0xFFFFFu);
                            col[d] = (int)(raw[d] >> 20);
                        }
                        if (q0 + PF < lk) {
#pragma unroll
                            for (int d = 0; d < PF; d++)
                                raw[d] = A.ntab_ent[off[k] + (int64_t)min(q0 + PF + d, lk - 1) * 64];
                        }
#pragma unroll
                        for (int r = 0; r < RG; r++)
#pragma unroll
                            for (int g = 0; g < NG; g++) {
                                if (rb[r] + ch * PCH + g * 4 >= rb[r + 1]) continue;   // uniform: group not staged
                                const int so = r * PCH + g * 4;
                                T4 vv[PF][NC], xv[PF][NV];
#pragma unroll
                                for (int d = 0; d < PF; d++) {
                                    if (q0 + d >= lk) continue;                  // wave-uniform
#pragma unroll
                                    for (int c = 0; c < NC; c++)
                                        vv[d][c] = *reinterpret_cast<const T4*>(
                                            Vs + ((size_t)c * A.cols2_max + col[d]) * SL + so);
#pragma unroll
                                    for (int b = 0; b < NV; b++)
                                        xv[d][b] = *reinterpret_cast<const T4*>(X + ((size_t)in[d] * NV + b) * SL + so);
                                }
#pragma unroll
                                for (int d = 0; d < PF; d++) {
                                    if (q0 + d >= lk) continue;                  // wave-uniform
#pragma unroll
                                    for (int b = 0; b < NV; b++) {
                                        const T4 v4 = vv[d][PV ? b : 0];
                                        T a = acc[k][r][b];
                                        a += v4.x * xv[d][b].x;
                                        a += v4.y * xv[d][b].y;
                                        a += v4.z * xv[d][b].z;
                                        a += v4.w * xv[d][b].w;
                                        acc[k][r][b] = a;
                                    }
                                }
                            }
                    }
                }
            } else {
                // Wide blocks (NPT > 4), all columns per entry step (measured on Mn50
                // FP64: 13.2 ms against 14.3 ms column by column).
                // Neutron-hop entries in chunks of PF steps, double-buffered: a chunk is
                // decoded into registers, the next chunk's loads are issued into the
                // raw buffer, and the decoded chunk is computed while they are in
                // flight.  Loads are unconditional (a column past its slice length
                // re-reads its last entry; every slice has at least one); the LDS reads
                // and FMAs of a column stop at its slice length (wave-uniform).
                constexpr int PF = NPT <= 4 ? 4 : (NPT <= 8 ? 2 : 1);
                uint32_t raw[PF][NPT];
#pragma unroll
                for (int d = 0; d < PF; d++)
#pragma unroll
                    for (int k = 0; k < NPT; k++)
                        raw[d][k] = A.ntab_ent[off[k] + (int64_t)min(d, max(len[k] - 1, 0)) * 64];
                for (int q0 = 0; q0 < maxlen; q0 += PF) {
                    int in[PF][NPT], col[PF][NPT];
#pragma unroll
                    for (int d = 0; d < PF; d++)
#pragma unroll
                        for (int k = 0; k < NPT; k++) {
                            in[d][k] = (int)(raw[d][k] & 0xFFFFFu);
                            col[d][k] = (int)(raw[d][k] >> 20);
                        }
                    if (q0 + PF < maxlen) {
#pragma unroll
                        for (int d = 0; d < PF; d++)
#pragma unroll
                            for (int k = 0; k < NPT; k++)
                                raw[d][k] = A.ntab_ent[off[k] + (int64_t)min(q0 + PF + d, max(len[k] - 1, 0)) * 64];
                    }
#pragma unroll
                    for (int d = 0; d < PF; d++) {
                        const int qq = q0 + d;
#pragma unroll
                        for (int r = 0; r < RG; r++)
#pragma unroll
                            for (int g = 0; g < NG; g++) {
                                if (rb[r] + ch * PCH + g * 4 >= rb[r + 1]) continue;   // uniform: group not staged
                                const int so = r * PCH + g * 4;
#pragma unroll
                                for (int k = 0; k < NPT; k++) {
                                    if (qq >= len[k]) continue;          // wave-uniform
                                    T4 vv = *reinterpret_cast<const T4*>(Vs + (size_t)col[d][k] * SL + so);
#pragma unroll
                                    for (int b = 0; b < NV; b++) {
                                        if (PV && b > 0)
                                            vv = *reinterpret_cast<const T4*>(Vs + ((size_t)b * A.cols2_max + col[d][k]) * SL + so);
                                        const T4 xv =
                                            *reinterpret_cast<const T4*>(X + ((size_t)in[d][k] * NV + b) * SL + so);
                                        T a = acc[k][r][b];
                                        a += vv.x * xv.x;
                                        a += vv.y * xv.y;
                                        a += vv.z * xv.z;
                                        a += vv.w * xv.w;
                                        acc[k][r][b] = a;
                                    }
                                }
                            }
                    }
                }
            }
        }
    }
    // The pn sums were accumulated in the block's walk order (lane slot j of column k
    // holds neutron SD nperm[j]); move them to the natural order through LDS (the X
    // region, free after the last round) before the diagonal, nn and pp terms.
    if (q1 > A.pair_ptr[bf]) {
        __syncthreads();
        T* Px = X;
        const int32_t* pm = A.nperm + A.en_off[bf];
#pragma unroll
        for (int k = 0; k < NPT; k++) {
            const int j = tid + k * TPB;
            if (j < nN) {
                const int f = pm[j];
#pragma unroll
                for (int r = 0; r < RG; r++)
#pragma unroll
                    for (int b = 0; b < NV; b++) Px[(f * RG + r) * NV + b] = acc[k][r][b];
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < NPT; k++) {
            const int j = tid + k * TPB;
#pragma unroll
            for (int r = 0; r < RG; r++)
#pragma unroll
                for (int b = 0; b < NV; b++) acc[k][r][b] = j < nN ? Px[(j * RG + r) * NV + b] : T(0);
        }
    }
    // diagonal (e_p + e_n) and nn (sliced ELL, entries loaded four at a time)
#pragma unroll
    for (int k = 0; k < NPT; k++) {
        const int sk = wave + NW * k;
        if (sk < nsl) {
            const int fnc = min(tid + k * TPB, nN - 1);
            T e_n[NC];
#pragma unroll
            for (int c = 0; c < NC; c++) e_n[c] = cen(c)[fnc];
#pragma unroll
            for (int r = 0; r < RG; r++)
#pragma unroll
                for (int b = 0; b < NV; b++)
                    acc[k][r][b] += (epr[r][PV ? b : 0] + e_n[PV ? b : 0]) * xs[(fnc * RG + r) * NV + b];
            const int64_t s = nsb + sk;
            const int64_t off = A.nn_slice_off[s] + (tid & 63);
            const int len = __builtin_amdgcn_readfirstlane(A.nn_slice_len[s]);
            int q = 0;
            for (; q + 4 <= len; q += 4) {
                T c[4][NC];
                int col[4];
#pragma unroll
                for (int u = 0; u < 4; u++) {
#pragma unroll
                    for (int cc = 0; cc < NC; cc++) c[u][cc] = cnn(cc)[off + (int64_t)(q + u) * 64];
                    col[u] = A.nn_col[off + (int64_t)(q + u) * 64];
                }
#pragma unroll
                for (int u = 0; u < 4; u++)
#pragma unroll
                    for (int r = 0; r < RG; r++)
#pragma unroll
                        for (int b = 0; b < NV; b++)
                            acc[k][r][b] += c[u][PV ? b : 0] * xs[(col[u] * RG + r) * NV + b];
            }
            for (; q < len; q++) {
                T c[NC];
#pragma unroll
                for (int cc = 0; cc < NC; cc++) c[cc] = cnn(cc)[off + (int64_t)q * 64];
                const int col = A.nn_col[off + (int64_t)q * 64];
#pragma unroll
                for (int r = 0; r < RG; r++)
#pragma unroll
                    for (int b = 0; b < NV; b++) acc[k][r][b] += c[PV ? b : 0] * xs[(col * RG + r) * NV + b];
            }
        }
    }
    // pp: rows of the same block; UP entries (uniform) per step, their row gathers
    // issued together as unconditional loads
    {
        constexpr int UP = NPT * NV <= 2 ? 8 : (NPT * NV <= 4 ? 4 : 2);
#pragma unroll
        for (int r = 0; r < RG; r++) {
            if (r < nr) {
                const int64_t pe0 = A.pp_rowptr[row0 + r], pe1 = A.pp_rowptr[row0 + r + 1];
                for (int64_t e = pe0; e < pe1; e += UP) {
                    T c[UP][NC];
                    int64_t so[UP];
#pragma unroll
                    for (int u = 0; u < UP; u++) {
                        const int64_t eu = min(e + u, pe1 - 1);
#pragma unroll
                        for (int cc = 0; cc < NC; cc++) c[u][cc] = e + u < pe1 ? cpp(cc)[eu] : T(0);
                        so[u] = bstart + (int64_t)A.pp_col[eu] * nN;
                    }
#pragma unroll
                    for (int k = 0; k < NPT; k++) {
                        if (wave + NW * k >= nsl) continue;         // wave-uniform
                        const int fnc = min(tid + k * TPB, nN - 1);
#pragma unroll
                        for (int b = 0; b < NV; b++) {
                            T x[UP];
#pragma unroll
                            for (int u = 0; u < UP; u++) x[u] = vin[b][so[u] + fnc];
                            T a = acc[k][r][b];
#pragma unroll
                            for (int u = 0; u < UP; u++) a += c[u][PV ? b : 0] * x[u];
                            acc[k][r][b] = a;
                        }
                    }
                }
            }
        }
    }
#pragma unroll
    for (int r = 0; r < RG; r++) {
        if (r < nr) {
#pragma unroll
            for (int b = 0; b < NV; b++) {
                T* dst = static_cast<T*>(P.w[b]) + bstart + (int64_t)(fp0 + r) * nN;
#pragma unroll
                for (int k = 0; k < NPT; k++) {
                    const int fn = tid + k * TPB;
                    if (fn < nN) dst[fn] = acc[k][r][b];
                }
            }
        }
    }
}

// The dynamic-LDS limit is a property of the kernel function, shared by every
// index that launches it: only ever raise it (indices of different widths may
// be resident and in use from several threads at once).
static void raise_lds_limit(const void* fn, size_t bytes) {
    static std::mutex mu;
    static std::map<const void*, size_t> set;
    std::lock_guard<std::mutex> lk(mu);
    size_t& cur = set[fn];
    if (bytes <= cur) return;
    SM_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    cur = bytes;
}

// instantiated shapes (NPT, RG, PCH, NV, PV, TPB): 256-thread workgroups for every
// width, one vector, two and four vectors on one Hamiltonian (four up to 1,536
// neutron determinants per block), two interaction variants on one index (PV);
// 512- and 1,024-thread workgroups (one row each) for wide neutron blocks, where
// the staged input rows of one workgroup fill most of a CU's LDS and more waves
// then share them
#define HV_NPTS(X_, T_, R_, P_, V_, W_) X_(T_, 1, R_, P_, V_, W_, 256) X_(T_, 2, R_, P_, V_, W_, 256) \
    X_(T_, 3, R_, P_, V_, W_, 256) X_(T_, 4, R_, P_, V_, W_, 256) X_(T_, 6, R_, P_, V_, W_, 256) \
    X_(T_, 8, R_, P_, V_, W_, 256) X_(T_, 12, R_, P_, V_, W_, 256) X_(T_, 16, R_, P_, V_, W_, 256)
#define HV_NPTS_NARROW(X_, T_, R_, P_, V_, W_) X_(T_, 1, R_, P_, V_, W_, 256) X_(T_, 2, R_, P_, V_, W_, 256) \
    X_(T_, 3, R_, P_, V_, W_, 256) X_(T_, 4, R_, P_, V_, W_, 256) X_(T_, 6, R_, P_, V_, W_, 256)
#define HV_NPTS_WIDE(X_, T_, V_, W_) X_(T_, 2, 1, 4, V_, W_, 512) X_(T_, 3, 1, 4, V_, W_, 512) \
    X_(T_, 4, 1, 4, V_, W_, 512) X_(T_, 6, 1, 4, V_, W_, 512) X_(T_, 8, 1, 4, V_, W_, 512) \
    X_(T_, 2, 1, 4, V_, W_, 1024) X_(T_, 3, 1, 4, V_, W_, 1024) X_(T_, 4, 1, 4, V_, W_, 1024)
#define HV_FOR_SHAPES(X_, T_) HV_NPTS(X_, T_, 2, 4, 1, false) HV_NPTS(X_, T_, 1, 4, 1, false) \
    HV_NPTS(X_, T_, 2, 4, 2, false) HV_NPTS(X_, T_, 1, 4, 2, false) \
    HV_NPTS_NARROW(X_, T_, 2, 4, 4, false) HV_NPTS_NARROW(X_, T_, 1, 4, 4, false) \
    HV_NPTS(X_, T_, 2, 4, 2, true) HV_NPTS(X_, T_, 1, 4, 2, true) \
    HV_NPTS_WIDE(X_, T_, 1, false) HV_NPTS_WIDE(X_, T_, 2, false) HV_NPTS_WIDE(X_, T_, 2, true)

template <typename T> static const void* hv_fn(int npt, int rg, int pch, int nv, bool pv, int tpb) {
#define HV_CASE(T_, N_, R_, P_, V_, W_, B_)                                                       \
    if (npt == N_ && rg == R_ && pch == P_ && nv == V_ && pv == W_ && tpb == B_)                   \
        return (const void*)k_hv<T_, N_, R_, P_, V_, W_, B_>;
    HV_FOR_SHAPES(HV_CASE, T)
#undef HV_CASE
    return nullptr;
}

// LDS of one workgroup for nv vectors: X (nN_max x slots x nv), Vs (cols2_max x
// slots, per vector for variants), own rows (nN_max x rg x nv)
static size_t hv_lds_bytes(const HvTables& A, int nv, bool pv, size_t tsize) {
    return ((size_t)A.nN_max * A.pch * A.rg * nv + (size_t)A.cols2_max * A.pch * A.rg * (pv ? nv : 1) +
            (size_t)A.nN_max * A.rg * nv) * tsize;
}

// Widest instantiated vector count <= nv whose workgroup LDS stays within 32 KiB, and
// for FP32 on 512-thread shapes two vectors within 80 KiB (two workgroups per CU, the
// occupancy the one-vector launch has there).  Measured per-vector times (FP32 / FP64):
// Cr48 (14.5 KiB per vector) 0.49 -> 0.34 ms with two vectors and 0.36 with four
// (48 KiB); Mg24 0.044 -> 0.017 ms with four.  Five-neutron pf blocks (1,651 wide,
// 512 threads, about 66 KiB at two vectors), FP32, one vector -> two of one
// Hamiltonian / two interaction variants: Cr49 1.53 -> 1.07 / 1.24 ms, Mn50 5.30 ->
// 4.05 / 4.58, Fe51 14.95 -> 11.74 / 13.08.  Slower: the same in FP64 (Mn50 8.66 ->
// 10.97, 132 KiB), six-neutron blocks (3,952 wide, 158 KiB: Fe52 39.4 -> 43.9, Mn51
// 13.8 -> 15.3), Fe50 FP64 at two (5.91 -> 6.66).  Results are bit-identical either way.
static constexpr size_t HV_LDS_SHARED = 32 * 1024;
static constexpr size_t HV_LDS_SHARED_WIDE = 80 * 1024;
template <typename T> static int hv_width(const HvTables& A, int nv, bool pv) {
    for (int c : {4, 2}) {
        if (c > nv || !hv_fn<T>(A.npt, A.rg, A.pch, c, pv, A.tpb)) continue;
        const size_t lds = hv_lds_bytes(A, c, pv, sizeof(T));
        if (lds <= HV_LDS_SHARED) return c;
        if (c == 2 && sizeof(T) == 4 && A.tpb == 512 && lds <= HV_LDS_SHARED_WIDE) return c;
    }
    return 1;
}

// Two vectors of one Hamiltonian where the basis' own 512-thread FP32 shape cannot take
// them (six-neutron pf blocks, 3,952 wide: 158 KiB at two vectors): the same tables at
// 1024 threads (columns fn = tid + 1024 k, the same 64-lane slices), one 16-wave
// workgroup per CU, the waves per CU of the one-vector launch (72 VGPRs).  Measured per
// vector, bit-identical: Mn51 13.79 -> 11.74 ms, Fe52 39.42 -> 33.20.  Not for
// interaction variants: their per-vector V_eff reads cancel the gain (Mn51 13.90, Fe52
// 38.80).
template <typename T> static bool wide_pair(const HvTables& A, HvTables& B) {
    if (sizeof(T) != 4 || A.tpb != 512) return false;
    B = A;
    B.tpb = 1024;
    B.npt = (A.nN_max + 1023) / 1024;
    return hv_fn<T>(B.npt, B.rg, B.pch, 2, false, 1024) && hv_lds_bytes(B, 2, false, sizeof(T)) <= 160 * 1024;
}

template <typename T>
static void launch_t(const HvTables& A, int nv, bool pv, const HvVecs& P, hipStream_t s) {
    dim3 grid(A.G), block(A.tpb);
    if (nv == 1) pv = false;
    const size_t lds = hv_lds_bytes(A, nv, pv, sizeof(T));
    if (nv > 1) raise_lds_limit(hv_fn<T>(A.npt, A.rg, A.pch, nv, pv, A.tpb), lds);
    bool done = false;
#define HV_LAUNCH(T_, N_, R_, P_, V_, W_, B_)                                                                \
    if (!done && A.npt == N_ && A.rg == R_ && A.pch == P_ && nv == V_ && pv == W_ && A.tpb == B_) {          \
        hipLaunchKernelGGL((k_hv<T_, N_, R_, P_, V_, W_, B_>), grid, block, lds, s, A, P);                  \
        done = true;                                                                                         \
    }
    HV_FOR_SHAPES(HV_LAUNCH, T)
#undef HV_LAUNCH
    if (!done) throw std::runtime_error("no H·v kernel for this neutron block width");
    SM_HIP(hipGetLastError());
}

void launch_hv(const DeviceHam& h, int prec, const void* v, void* w, hipStream_t s) {
    if (h.idx->R == 0) return;
    HvVecs P{};
    P.v[0] = v;
    P.w[0] = w;
    if (prec == 0) launch_t<double>(h.view[0], 1, false, P, s);
    else launch_t<float>(h.view[1], 1, false, P, s);
}

void launch_hv_column(const DeviceHam& h, int prec, const void* v, int64_t ld, const int* vsel, void* w, hipStream_t s) {
    if (h.idx->R == 0) return;
    HvVecs P{};
    P.v[0] = v;
    P.w[0] = w;
    P.vsel = vsel;
    P.vld = ld;
    if (prec == 0) launch_t<double>(h.view[0], 1, false, P, s);
    else launch_t<float>(h.view[1], 1, false, P, s);
}

template <typename T>
static void launch_set_t(const HvTables& A, int nvec, const void* const* v, void* const* w, hipStream_t s) {
    for (int b = 0; b < nvec;) {
        int c = hv_width<T>(A, nvec - b, false);
        HvTables B;
        const HvTables* L = &A;
        if (c == 1 && nvec - b >= 2 && wide_pair<T>(A, B)) {
            c = 2;
            L = &B;
        }
        HvVecs P{};
        for (int j = 0; j < c; j++) { P.v[j] = v[b + j]; P.w[j] = w[b + j]; }
        launch_t<T>(*L, c, false, P, s);
        b += c;
    }
}

void launch_hv_set(const DeviceHam& h, int prec, int nvec, const void* const* v, void* const* w, hipStream_t s) {
    if (h.idx->R == 0 || nvec <= 0) return;
    if (prec == 0) launch_set_t<double>(h.view[0], nvec, v, w, s);
    else launch_set_t<float>(h.view[1], nvec, v, w, s);
}

// Runs of one Hamiltonian (same handle) go to the shared-coefficient kernels (up to
// four vectors); the rest are paired into per-vector-coefficient launches.
template <typename T>
static void launch_variants_t(const DeviceHam* const* hams, int prec, int nvec, const void* const* v,
                              void* const* w, hipStream_t s) {
    const HvTables& A = hams[0]->view[prec];
    for (int b = 0; b < nvec;) {
        int run = 1;
        while (b + run < nvec && hams[b + run] == hams[b]) run++;
        if (run >= 2) {
            int c = hv_width<T>(A, run, false);
            HvTables B;
            if (c == 1 && wide_pair<T>(A, B)) c = 2;
            launch_set_t<T>(hams[b]->view[prec], c, v + b, w + b, s);
            b += c;
            continue;
        }
        const int c = hv_width<T>(A, nvec - b, true);
        HvVecs P{};
        for (int j = 0; j < c; j++) {
            const HvTables& Aj = hams[b + j]->view[prec];
            P.v[j] = v[b + j];
            P.w[j] = w[b + j];
            P.ep[j] = Aj.ep;
            P.en[j] = Aj.en;
            P.pp[j] = Aj.pp_val;
            P.nn[j] = Aj.nn_val;
            P.veff[j] = Aj.veff;
        }
        launch_t<T>(c == 1 ? hams[b]->view[prec] : A, c, c > 1, P, s);
        b += c;
    }
}

void launch_hv_variants(const DeviceHam* const* hams, int prec, int nvec, const void* const* v, void* const* w,
                        hipStream_t s) {
    if (nvec <= 0 || hams[0]->idx->R == 0) return;
    for (int b = 1; b < nvec; b++)
        if (hams[b]->idx != hams[0]->idx) throw std::invalid_argument("variants must share one index");
    if (prec == 0) launch_variants_t<double>(hams, prec, nvec, v, w, s);
    else launch_variants_t<float>(hams, prec, nvec, v, w, s);
}

void launch_hv_multi(const DeviceHam& h, int prec, int nvec, const void* v, int64_t ldv, void* w, int64_t ldw,
                     hipStream_t s) {
    if (h.idx->R == 0 || nvec <= 0) return;
    const size_t es = prec == 0 ? 8 : 4;
    std::vector<const void*> vp(nvec);
    std::vector<void*> wp(nvec);
    for (int b = 0; b < nvec; b++) {
        vp[b] = static_cast<const char*>(v) + (size_t)b * ldv * es;
        wp[b] = static_cast<char*>(w) + (size_t)b * ldw * es;
    }
    launch_hv_set(h, prec, nvec, vp.data(), wp.data(), s);
}

// ---------------------------------------------------------------- upload
namespace {

template <typename T> std::vector<T> cast_vec(const std::vector<double>& v) {
    std::vector<T> o(v.size());
    for (size_t i = 0; i < v.size(); i++) o[i] = (T)v[i];
    return o;
}

struct HostLayout {
    std::vector<int32_t> row_block, block_row0, nP, nN;
    std::vector<int64_t> tdims, en_off;
    std::vector<double> ep, en;
    std::vector<int64_t> pp_rowptr;
    std::vector<int32_t> pp_col;
    std::vector<double> pp_val;
    std::vector<int64_t> nn_slice_base, nn_slice_off;
    std::vector<int32_t> nn_slice_len, nn_col;
    std::vector<double> nn_val;
    std::vector<int32_t> pair_ptr, pair_bi, pair_dm, pair_ptab, pair_ntab;
    std::vector<int64_t> ptab_rowptr_base, ptab_ent_base;
    std::vector<int32_t> ptab_rowptr;
    std::vector<uint32_t> ptab_ent;
    std::vector<int64_t> ntab_slice_base, ntab_slice_off;
    std::vector<int32_t> ntab_slice_len;
    std::vector<int32_t> nperm;        // per block (offset en_off): pn lane slot -> neutron SD
    std::vector<uint32_t> ntab_ent;
    std::vector<int64_t> veff_off;
    std::vector<int32_t> veff_cols2;
    std::vector<double> veff;
    int nN_max = 0, cols2_max = 0;
    int64_t pn_terms = 0, pn_terms_exact = 0, pp_terms = 0, nn_terms = 0;
};

void coef_layout(const Basis& b, const Tables& t, HostLayout& L) {
    L.ep.clear(); L.en.clear(); L.pp_val.clear(); L.nn_val.clear(); L.veff.clear();
    for (int k = 0; k < b.nb; k++) {
        const auto& e = t.ep[b.pset[k]];
        L.ep.insert(L.ep.end(), e.begin(), e.end());
        const auto& f = t.en[b.nset[k]];
        L.en.insert(L.en.end(), f.begin(), f.end());
        const PairSym& m = t.pp[b.pset[k]];
        L.pp_val.insert(L.pp_val.end(), m.val.begin(), m.val.end());
        const PairSym& nm = t.nn[b.nset[k]];
        int nN = (int)b.sp[1].sets[b.nset[k]].size();
        for (int s0 = 0; s0 < nN; s0 += 64) {
            int len = 0;
            for (int f = s0; f < std::min(nN, s0 + 64); f++) len = std::max<int>(len, (int)(nm.rowptr[f + 1] - nm.rowptr[f]));
            for (int q = 0; q < len; q++)
                for (int lane = 0; lane < 64; lane++) {
                    int f = s0 + lane;
                    double val = 0.0;
                    if (f < nN && nm.rowptr[f] + q < nm.rowptr[f + 1]) val = nm.val[nm.rowptr[f] + q];
                    L.nn_val.push_back(val);
                }
        }
    }
    for (size_t k = 0; k < t.veff.size(); k++) {
        int rows = t.veff_rows[k], cols = t.veff_cols[k];
        int c2 = 2 * (cols + 1);
        for (int r = 0; r < rows; r++) {
            for (int c = 0; c < cols; c++) {
                double v = t.veff[k][(size_t)r * cols + c];
                L.veff.push_back(v);
                L.veff.push_back(-v);
            }
            L.veff.push_back(0.0);
            L.veff.push_back(0.0);
        }
        (void)c2;
    }
}

void index_layout(const Basis& b, const Tables& t, HostLayout& L) {
    const Species& P = b.sp[0];
    const Species& N = b.sp[1];
    L.tdims = b.tdims;
    int R = 0;
    for (int k = 0; k < b.nb; k++) {
        int np = (int)P.sets[b.pset[k]].size(), nn = (int)N.sets[b.nset[k]].size();
        L.block_row0.push_back(R);
        L.nP.push_back(np);
        L.nN.push_back(nn);
        L.nN_max = std::max(L.nN_max, nn);
        for (int i = 0; i < np; i++) L.row_block.push_back(k);
        R += np;
    }
    // pp CSR per global row
    L.pp_rowptr.push_back(0);
    for (int k = 0; k < b.nb; k++) {
        const PairSym& m = t.pp[b.pset[k]];
        int np = L.nP[k];
        for (int i = 0; i < np; i++) {
            for (int64_t e = m.rowptr[i]; e < m.rowptr[i + 1]; e++) L.pp_col.push_back(m.col[e]);
            L.pp_rowptr.push_back((int64_t)L.pp_col.size());
            L.pp_terms += (m.rowptr[i + 1] - m.rowptr[i]) * (int64_t)L.nN[k];
        }
    }
    // nn sliced ELL per block
    int64_t en_acc = 0;
    for (int k = 0; k < b.nb; k++) {
        L.en_off.push_back(en_acc);
        en_acc += L.nN[k];
        const PairSym& m = t.nn[b.nset[k]];
        int nN = L.nN[k];
        L.nn_slice_base.push_back((int64_t)L.nn_slice_off.size());
        for (int s0 = 0; s0 < nN; s0 += 64) {
            int len = 0;
            for (int f = s0; f < std::min(nN, s0 + 64); f++) len = std::max<int>(len, (int)(m.rowptr[f + 1] - m.rowptr[f]));
            L.nn_slice_off.push_back((int64_t)L.nn_col.size());
            L.nn_slice_len.push_back(len);
            for (int q = 0; q < len; q++)
                for (int lane = 0; lane < 64; lane++) {
                    int f = s0 + lane;
                    int col = 0;
                    if (f < nN && m.rowptr[f] + q < m.rowptr[f + 1]) col = m.col[m.rowptr[f] + q];
                    L.nn_col.push_back(col);
                }
        }
        for (int f = 0; f < nN; f++) L.nn_terms += (m.rowptr[f + 1] - m.rowptr[f]) * (int64_t)L.nP[k];
    }
    // The pn walk visits a block's neutron SDs in an order of its own: sorted by their
    // walk cost summed over the block's input blocks (entries x staging rounds), so
    // the 64 SDs of a slice have similar entry counts and the wave-uniform walk to the
    // slice's longest row pads less.  Measured on the tables: padded / exact pn terms
    // Cr48 1.352 -> 1.232, Mn50 1.292 -> 1.137, Fe52 1.266 -> 1.091.  The kernel
    // accumulates in this order and moves the sums to the natural order once.  Measured
    // against the natural order (ms per launch FP32 / FP64): Cr48 0.460 -> 0.450 /
    // 0.680 -> 0.671, Mn50 5.49 -> 5.32 / 9.13 -> 8.77, Fe52 40.5 -> 39.4 / 78.1 -> 76.0;
    // sorting only within windows of 128-512 SDs gained less.
    std::vector<int> block_of_nset(b.sp[1].sets.size(), -1);
    for (int k = 0; k < b.nb; k++) block_of_nset[b.nset[k]] = k;
    std::vector<std::vector<int32_t>> perm(b.nb);
    {
        std::vector<std::vector<double>> key(b.nb);
        for (int k = 0; k < b.nb; k++) key[k].assign(L.nN[k], 0.0);
        for (int bf = 0; bf < b.nb; bf++)
            for (int bi = 0; bi < b.nb; bi++) {
                const int dM = b.Mp[bf] - b.Mp[bi];
                if (t.ob[0].index_of_dM(dM) < 0 || t.ob[1].index_of_dM(-dM) < 0) continue;
                auto ip = t.hop_index.find({0, b.pset[bi], b.pset[bf]});
                auto in = t.hop_index.find({1, b.nset[bi], b.nset[bf]});
                if (ip == t.hop_index.end() || in == t.hop_index.end()) continue;
                const HopTable& HP = t.hops[ip->second];
                const HopTable& HN = t.hops[in->second];
                double rounds = 0;
                for (size_t r = 0; r + 1 < HP.rowptr.size(); r++) rounds += (double)((HP.rowptr[r + 1] - HP.rowptr[r] + 3) / 4);
                for (int f = 0; f < L.nN[bf]; f++) key[bf][f] += rounds * (double)(HN.rowptr[f + 1] - HN.rowptr[f]);
            }
        for (int k = 0; k < b.nb; k++) {
            perm[k].resize(L.nN[k]);
            for (int f = 0; f < L.nN[k]; f++) perm[k][f] = f;
            std::stable_sort(perm[k].begin(), perm[k].end(), [&](int a, int c) { return key[k][a] > key[k][c]; });
            L.nperm.insert(L.nperm.end(), perm[k].begin(), perm[k].end());
        }
    }
    // hop tables: proton tables keep CSR, neutron tables become sliced ELL (rows in
    // the output block's walk order)
    std::vector<int> pt_of(t.hops.size(), -1), nt_of(t.hops.size(), -1);
    for (size_t h = 0; h < t.hops.size(); h++) {
        const HopTable& H = t.hops[h];
        if (H.species == 0) {
            pt_of[h] = (int)L.ptab_rowptr_base.size();
            L.ptab_rowptr_base.push_back((int64_t)L.ptab_rowptr.size());
            L.ptab_ent_base.push_back((int64_t)L.ptab_ent.size());
            for (int64_t x : H.rowptr) L.ptab_rowptr.push_back((int32_t)x);
            L.ptab_ent.insert(L.ptab_ent.end(), H.ent.begin(), H.ent.end());
        } else {
            nt_of[h] = (int)L.ntab_slice_base.size();
            L.ntab_slice_base.push_back((int64_t)L.ntab_slice_off.size());
            int kp = t.ob[0].index_of_dM(-H.dM);
            int cols = kp >= 0 ? t.veff_cols[kp] : 0;
            int nf = (int)H.rowptr.size() - 1;
            const int bfk = block_of_nset[H.fset];
            if (bfk < 0 || (int)perm[bfk].size() != nf) throw std::logic_error("neutron table without its block");
            const std::vector<int32_t>& pm = perm[bfk];
            for (int s0 = 0; s0 < nf; s0 += 64) {
                // at least one row per slice (zero-column entries): the kernel reads a
                // slice's last row unconditionally when prefetching past its length
                int len = 1;
                for (int j = s0; j < std::min(nf, s0 + 64); j++)
                    len = std::max<int>(len, (int)(H.rowptr[pm[j] + 1] - H.rowptr[pm[j]]));
                L.ntab_slice_off.push_back((int64_t)L.ntab_ent.size());
                L.ntab_slice_len.push_back(len);
                for (int q = 0; q < len; q++)
                    for (int lane = 0; lane < 64; lane++) {
                        const int j = s0 + lane;
                        uint32_t ent = (uint32_t)(2 * cols) << 20;         // zero column
                        const int f = j < nf ? pm[j] : 0;
                        if (j < nf && H.rowptr[f] + q < H.rowptr[f + 1]) {
                            uint32_t e = H.ent[H.rowptr[f] + q];
                            uint32_t in = e & 0xFFFFFu, op = (e >> 20) & 0x7FFu, neg = e >> 31;
                            ent = in | ((2 * op + neg) << 20);
                        }
                        L.ntab_ent.push_back(ent);
                    }
            }
        }
    }
    // V_eff offsets
    int64_t voff = 0;
    for (size_t k = 0; k < t.veff.size(); k++) {
        int c2 = 2 * (t.veff_cols[k] + 1);
        if (c2 >= 4096) throw std::runtime_error("too many neutron one-body operators for the packed entry");
        L.veff_off.push_back(voff);
        L.veff_cols2.push_back(c2);
        L.cols2_max = std::max(L.cols2_max, c2);
        voff += (int64_t)t.veff_rows[k] * c2;
    }
    // block pairs (output bf, input bi)
    L.pair_ptr.push_back(0);
    for (int bf = 0; bf < b.nb; bf++) {
        for (int bi = 0; bi < b.nb; bi++) {
            int dM = b.Mp[bf] - b.Mp[bi];
            int k = t.ob[0].index_of_dM(dM);
            if (k < 0 || t.ob[1].index_of_dM(-dM) < 0) continue;
            auto ip = t.hop_index.find({0, b.pset[bi], b.pset[bf]});
            auto in = t.hop_index.find({1, b.nset[bi], b.nset[bf]});
            if (ip == t.hop_index.end() || in == t.hop_index.end()) continue;
            L.pair_bi.push_back(bi);
            L.pair_dm.push_back(k);
            L.pair_ptab.push_back(pt_of[ip->second]);
            L.pair_ntab.push_back(nt_of[in->second]);
            const HopTable& HP = t.hops[ip->second];
            const HopTable& HN = t.hops[in->second];
            int64_t nsum = 0;
            for (size_t f = 0; f + 1 < HN.rowptr.size(); f++) nsum += HN.rowptr[f + 1] - HN.rowptr[f];
            int64_t npad = 0;
            int nf = (int)HN.rowptr.size() - 1;
            const std::vector<int32_t>& pm = perm[bf];
            for (int s0 = 0; s0 < nf; s0 += 64) {
                int len = 1;
                for (int j = s0; j < std::min(nf, s0 + 64); j++)
                    len = std::max<int>(len, (int)(HN.rowptr[pm[j] + 1] - HN.rowptr[pm[j]]));
                npad += (int64_t)len * std::min(64, nf - s0);
            }
            int64_t psum = (int64_t)HP.ent.size();
            L.pn_terms_exact += psum * nsum;
            L.pn_terms += psum * npad;
        }
        L.pair_ptr.push_back((int32_t)L.pair_bi.size());
    }
}


// LDS elements one one-vector k_hv launch moves (RG rows per workgroup, four hop
// slots per round; bytes = elements x the element size): [0] own rows staged and
// the pn sums moved from walk order to natural order, [1]
// diagonal reads, [2] nn reads, [3] pn staging writes (X rows and V_eff rows of the
// hops), [4] pn walk reads (per neutron-entry step of an active wave, one four-slot
// read of X and one of V_eff per staged row; lanes past the block edge included).
void lds_model(const HostLayout& L, int rg, const std::vector<int32_t>& gb, const std::vector<int32_t>& gf,
               int64_t* e) {
    for (int k = 0; k < 5; k++) e[k] = 0;
    const int nb = (int)L.nP.size();
    std::vector<int64_t> nn_len(nb, 0), ntab_len(L.ntab_slice_base.size(), 0);
    for (int b = 0; b < nb; b++) {
        const int nsl = (L.nN[b] + 63) / 64;
        for (int s = 0; s < nsl; s++) nn_len[b] += L.nn_slice_len[L.nn_slice_base[b] + s];
    }
    for (size_t q = 0; q < L.pair_bi.size(); q++) {
        const int nt = L.pair_ntab[q];
        if (ntab_len[nt]) continue;
        int bf = 0;
        while (L.pair_ptr[bf + 1] <= (int)q) bf++;
        const int nsl = (L.nN[bf] + 63) / 64;
        for (int s = 0; s < nsl; s++) ntab_len[nt] += L.ntab_slice_len[L.ntab_slice_base[nt] + s];
    }
    for (size_t g = 0; g < gb.size(); g++) {
        const int bf = gb[g], fp0 = gf[g];
        const int nN = L.nN[bf], nsl = (nN + 63) / 64;
        const int nr = std::min(rg, L.nP[bf] - fp0);
        e[0] += (int64_t)nN * rg * (L.pair_ptr[bf + 1] > L.pair_ptr[bf] ? 3 : 1);
        e[1] += (int64_t)64 * nsl * rg;
        e[2] += nn_len[bf] * 64 * rg;
        for (int q = L.pair_ptr[bf]; q < L.pair_ptr[bf + 1]; q++) {
            const int pt = L.pair_ptab[q];
            const int32_t* prp = L.ptab_rowptr.data() + L.ptab_rowptr_base[pt] + fp0;
            const int nNi = L.nN[L.pair_bi[q]], c2 = L.veff_cols2[L.pair_dm[q]];
            const int64_t walk = ntab_len[L.pair_ntab[q]] * 64;
            int nch = 0;
            for (int r = 0; r < nr; r++) nch = std::max(nch, (prp[r + 1] - prp[r] + 3) / 4);
            for (int ch = 0; ch < nch; ch++) {
                int staged = 0;
                for (int r = 0; r < nr; r++) staged += prp[r] + 4 * ch < prp[r + 1] ? 1 : 0;
                e[3] += (int64_t)staged * (nNi + c2) * 4;
                e[4] += (int64_t)staged * walk * 8;
            }
        }
    }
}

template <typename T> void set_view(DeviceIndex& h, int prec, const HostLayout& L) {
    HvTables& A = h.view[prec];
    A.nb = h.nb; A.R = h.R;
    A.row_block = h.row_block.as<int32_t>(); A.block_row0 = h.block_row0.as<int32_t>();
    A.nP = h.nP.as<int32_t>(); A.nN = h.nN.as<int32_t>(); A.tdims = h.tdims.as<int64_t>();
    A.ep = nullptr; A.en_off = h.en_off.as<int64_t>(); A.en = nullptr;
    A.pp_rowptr = h.pp_rowptr.as<int64_t>(); A.pp_col = h.pp_col.as<int32_t>(); A.pp_val = nullptr;
    A.nn_slice_base = h.nn_slice_base.as<int64_t>(); A.nn_slice_off = h.nn_slice_off.as<int64_t>();
    A.nn_slice_len = h.nn_slice_len.as<int32_t>(); A.nn_col = h.nn_col.as<int32_t>(); A.nn_val = nullptr;
    A.pair_ptr = h.pair_ptr.as<int32_t>(); A.pair_bi = h.pair_bi.as<int32_t>(); A.pair_dm = h.pair_dm.as<int32_t>();
    A.pair_ptab = h.pair_ptab.as<int32_t>(); A.pair_ntab = h.pair_ntab.as<int32_t>();
    A.ptab_rowptr_base = h.ptab_rowptr_base.as<int64_t>(); A.ptab_rowptr = h.ptab_rowptr.as<int32_t>();
    A.ptab_ent_base = h.ptab_ent_base.as<int64_t>(); A.ptab_ent = h.ptab_ent.as<uint32_t>();
    A.ntab_slice_base = h.ntab_slice_base.as<int64_t>(); A.ntab_slice_off = h.ntab_slice_off.as<int64_t>();
    A.ntab_slice_len = h.ntab_slice_len.as<int32_t>(); A.ntab_ent = h.ntab_ent.as<uint32_t>();
    A.nperm = h.nperm.as<int32_t>();
    A.veff_off = h.veff_off.as<int64_t>(); A.veff_cols2 = h.veff_cols2.as<int32_t>(); A.veff = nullptr;
    A.nN_max = std::max(1, L.nN_max);
    A.cols2_max = std::max(2, L.cols2_max);
    // LDS per workgroup: X (nN_max x slots), Vs (cols2_max x slots), own rows (nN_max x rg)
    auto lds_for = [&](int rg, int pch) {
        return ((size_t)A.nN_max * pch * rg + (size_t)A.cols2_max * pch * rg + (size_t)A.nN_max * rg) * sizeof(T);
    };
    // two-row groups for narrow blocks, one row otherwise; four slots per row and
    // round (measured on Cr48 and Mn50: 8 or 16 slots per round, i.e. fewer
    // staging rounds, are no faster and cost LDS occupancy)
    A.rg = lds_for(2, 4) <= HV_LDS_TWO_ROWS ? 2 : 1;
    A.pch = 4;
    const int npts[] = {1, 2, 3, 4, 6, 8, 12, 16};
    auto npt_for = [&](int tpb) {
        const int need = (A.nN_max + tpb - 1) / tpb;
        for (int x : npts) if (x >= need) return x;
        return 0;
    };
    // Workgroup size: blocks up to 1,024 neutron determinants run 256 threads (the
    // LDS of several workgroups fits a CU); wider blocks run one row per 512- or
    // 1,024-thread workgroup, whose 8 or 16 waves share one copy of the staged input
    // rows (with 256 threads the FP64 staging of a 3,952-wide block leaves one 4-wave
    // workgroup per CU).  Measured, ms per launch FP32 / FP64 at 256, 512, 1024
    // threads (results bit-identical): Mn50 (widest block 1,651) 5.49 5.42 6.92 /
    // 12.4 8.84 11.6; Mn51 (3,952) 22.5 14.0 14.8 / 50.1 33.9 27.8; Fe52 (3,952)
    // 65.0 40.0 41.8 / 144 96.0 76.7; Cr48 (565) 0.432 0.538 - / 0.708 0.816 -.
    // SMGPU_HV_TPB = 256 / 512 / 1024 overrides the choice where an instance exists.
    A.tpb = 256;
    if (A.nN_max > 1024) A.tpb = (sizeof(T) == 8 && A.nN_max > 2048) ? 1024 : 512;
    if (const char* e = std::getenv("SMGPU_HV_TPB")) {
        const int want = std::atoi(e);
        if (want == 256 || want == 512 || want == 1024) A.tpb = want;
    }
    if (A.tpb > 256) A.rg = 1;
    A.npt = npt_for(A.tpb);
    if (A.tpb > 256 && !hv_fn<T>(A.npt, 1, 4, 1, false, A.tpb)) {
        A.tpb = 256;
        A.rg = lds_for(2, 4) <= HV_LDS_TWO_ROWS ? 2 : 1;
        A.npt = npt_for(256);
    }
    if (A.npt == 0) throw std::runtime_error("neutron block wider than 4096 determinants");
    A.lds_bytes = lds_for(A.rg, A.pch);
    if (A.lds_bytes > 160 * 1024) throw std::runtime_error("neutron block too wide for the LDS of one CU");
    // Row groups in (block, row) order: the rows of one block stage many of the same
    // input rows.  Workgroup b runs on XCD b % 8, and each XCD has its own L2, so the
    // list is dealt to the XCDs in chunks of C consecutive groups (chunk j to XCD j % 8):
    // an XCD runs neighbouring rows of a block together.  The groups past the last full
    // deal keep their place.  Measured (ms per launch, FP32, Cr48 one / two vectors;
    // Mn50 one): longest-processing-time order (the reference's make_distribute)
    // 0.449 / 0.629; 6.01; dealt in chunks of 64: 0.438 / 0.610; 5.85; block order
    // dealt in chunks of 64: 0.411 / 0.572; 5.49.  C = 128 and 256 were slower on Cr48
    // (G = 4,845) with the cost order.
    std::vector<int32_t> gb, gf;
    for (int k = 0; k < h.nb; k++)
        for (int i = 0; i < L.nP[k]; i += A.rg) {
            gb.push_back(k);
            gf.push_back(i);
        }
    {
        const int G = (int)gb.size(), X = 8;
        const int C = std::max(1, std::min(64, G / 64));
        const int nfull = (G / (X * C)) * X * C;
        std::vector<int32_t> b2(gb), f2(gf);
        for (int b = 0; b < nfull; b++) {
            const int x = b % X, k = b / X;
            const int g = ((k / C) * X + x) * C + (k % C);
            b2[b] = gb[g];
            f2[b] = gf[g];
        }
        gb.swap(b2);
        gf.swap(f2);
    }
    h.grp_block[prec].upload(gb);
    h.grp_fp0[prec].upload(gf);
    lds_model(L, A.rg, gb, gf, h.lds_elems[prec]);
    A.G = (int)gb.size();
    A.grp_block = h.grp_block[prec].as<int32_t>();
    A.grp_fp0 = h.grp_fp0[prec].as<int32_t>();
    const void* fn = hv_fn<T>(A.npt, A.rg, A.pch, 1, false, A.tpb);
    if (!fn) throw std::runtime_error("no H·v kernel instance for this shape");
    raise_lds_limit(fn, A.lds_bytes);
}

}  // namespace

struct HostIndex {
    HostLayout L;          // index part
    uint64_t pattern = 0;
};

struct HostCoefficients {
    std::vector<double> ep, en, pp_val, nn_val, veff;
    std::vector<float> ep32, en32, pp_val32, nn_val32, veff32;
    uint64_t pattern = 0;
};

// hash of the FP64 coefficient values (the FP32 copies derive from them)
uint64_t content_hash(const HostCoefficients& c) {
    uint64_t hsh = 1469598103934665603ull;
    for (const std::vector<double>* v : {&c.ep, &c.en, &c.pp_val, &c.nn_val, &c.veff}) {
        hsh = (hsh ^ (uint64_t)v->size()) * 1099511628211ull;
        for (double x : *v) {
            uint64_t u;
            std::memcpy(&u, &x, 8);
            hsh = (hsh ^ u) * 1099511628211ull;
        }
    }
    return hsh;
}

bool same_coefficients(const HostCoefficients& a, const HostCoefficients& b) {
    return a.pattern == b.pattern && a.ep == b.ep && a.en == b.en && a.pp_val == b.pp_val && a.nn_val == b.nn_val &&
           a.veff == b.veff;
}

uint64_t coefficient_pattern(const Basis& b, const Tables& t) {
    uint64_t hsh = 1469598103934665603ull;
    auto mix = [&](uint64_t x) { hsh ^= x + 0x9E3779B97F4A7C15ull + (hsh << 6) + (hsh >> 2); };
    for (int k = 0; k < b.nb; k++) {
        for (const PairSym* m : {&t.pp[b.pset[k]], &t.nn[b.nset[k]]}) {
            mix(m->rowptr.size());
            for (int64_t x : m->rowptr) mix((uint64_t)x);
            for (int32_t c : m->col) mix((uint64_t)c);
        }
    }
    return hsh;
}

uint64_t pattern_of(const HostCoefficients& c) { return c.pattern; }
uint64_t pattern_of(const HostIndex& x) { return x.pattern; }

std::shared_ptr<const HostIndex> pack_index(const Basis& b, const Tables& t) {
    auto x = std::make_shared<HostIndex>();
    index_layout(b, t, x->L);
    x->pattern = coefficient_pattern(b, t);
    return x;
}

template <typename V> static int64_t vbytes(const V& v) { return (int64_t)(v.size() * sizeof(v[0])); }

int64_t device_bytes(const HostIndex& x) {
    const HostLayout& L = x.L;
    int64_t rows = (int64_t)L.row_block.size();
    // row groups of both precisions: at most one group per row, block + first row
    return vbytes(L.row_block) + vbytes(L.block_row0) + vbytes(L.nP) + vbytes(L.nN) + vbytes(L.tdims) +
           vbytes(L.en_off) + vbytes(L.pp_rowptr) + vbytes(L.pp_col) + vbytes(L.nn_slice_base) +
           vbytes(L.nn_slice_off) + vbytes(L.nn_slice_len) + vbytes(L.nn_col) + vbytes(L.pair_ptr) +
           vbytes(L.pair_bi) + vbytes(L.pair_dm) + vbytes(L.pair_ptab) + vbytes(L.pair_ntab) +
           vbytes(L.ptab_rowptr_base) + vbytes(L.ptab_rowptr) + vbytes(L.ptab_ent_base) + vbytes(L.ptab_ent) +
           vbytes(L.ntab_slice_base) + vbytes(L.ntab_slice_off) + vbytes(L.ntab_slice_len) + vbytes(L.ntab_ent) +
           vbytes(L.nperm) +
           vbytes(L.veff_off) + vbytes(L.veff_cols2) + 2 * 2 * rows * (int64_t)sizeof(int32_t);
}

std::shared_ptr<const HostCoefficients> pack_coefficients(const Basis& b, const Tables& t) {
    HostLayout L;
    coef_layout(b, t, L);
    auto c = std::make_shared<HostCoefficients>();
    c->ep32 = cast_vec<float>(L.ep); c->en32 = cast_vec<float>(L.en); c->pp_val32 = cast_vec<float>(L.pp_val);
    c->nn_val32 = cast_vec<float>(L.nn_val); c->veff32 = cast_vec<float>(L.veff);
    c->ep = std::move(L.ep); c->en = std::move(L.en); c->pp_val = std::move(L.pp_val);
    c->nn_val = std::move(L.nn_val); c->veff = std::move(L.veff);
    c->pattern = coefficient_pattern(b, t);
    return c;
}

int64_t device_bytes(const HostCoefficients& c) {
    return vbytes(c.ep) + vbytes(c.en) + vbytes(c.pp_val) + vbytes(c.nn_val) + vbytes(c.veff) + vbytes(c.ep32) +
           vbytes(c.en32) + vbytes(c.pp_val32) + vbytes(c.nn_val32) + vbytes(c.veff32);
}

int64_t coefficient_growth(const DeviceHam& h, const HostCoefficients& c) {
    int64_t g = 0;
    auto f = [&](const DBuf& d, int64_t need) { if (need > (int64_t)d.cap) g += need - (int64_t)d.cap; };
    f(h.ep[0], vbytes(c.ep)); f(h.en[0], vbytes(c.en)); f(h.pp_val[0], vbytes(c.pp_val));
    f(h.nn_val[0], vbytes(c.nn_val)); f(h.veff[0], vbytes(c.veff));
    f(h.ep[1], vbytes(c.ep32)); f(h.en[1], vbytes(c.en32)); f(h.pp_val[1], vbytes(c.pp_val32));
    f(h.nn_val[1], vbytes(c.nn_val32)); f(h.veff[1], vbytes(c.veff32));
    return g;
}

std::shared_ptr<DeviceIndex> upload_index(const HostIndex& x, int device) {
    SM_HIP(hipSetDevice(device));
    const HostLayout& L = x.L;
    auto h = std::make_shared<DeviceIndex>();
    h->device = device;
    h->mdim = L.tdims.empty() ? 0 : L.tdims.back();
    h->nb = (int)L.nP.size();
    h->R = (int)L.row_block.size();
    h->row_block.upload(L.row_block); h->block_row0.upload(L.block_row0);
    h->nP.upload(L.nP); h->nN.upload(L.nN); h->tdims.upload(L.tdims); h->en_off.upload(L.en_off);
    h->pp_rowptr.upload(L.pp_rowptr); h->pp_col.upload(L.pp_col);
    h->nn_slice_base.upload(L.nn_slice_base); h->nn_slice_off.upload(L.nn_slice_off);
    h->nn_slice_len.upload(L.nn_slice_len); h->nn_col.upload(L.nn_col);
    h->pair_ptr.upload(L.pair_ptr); h->pair_bi.upload(L.pair_bi); h->pair_dm.upload(L.pair_dm);
    h->pair_ptab.upload(L.pair_ptab); h->pair_ntab.upload(L.pair_ntab);
    h->ptab_rowptr_base.upload(L.ptab_rowptr_base); h->ptab_rowptr.upload(L.ptab_rowptr);
    h->ptab_ent_base.upload(L.ptab_ent_base); h->ptab_ent.upload(L.ptab_ent);
    h->ntab_slice_base.upload(L.ntab_slice_base); h->ntab_slice_off.upload(L.ntab_slice_off);
    h->ntab_slice_len.upload(L.ntab_slice_len); h->ntab_ent.upload(L.ntab_ent); h->nperm.upload(L.nperm);
    h->veff_off.upload(L.veff_off); h->veff_cols2.upload(L.veff_cols2);
    set_view<double>(*h, 0, L);
    set_view<float>(*h, 1, L);
    h->pattern = x.pattern;
    h->pn_terms = L.pn_terms;
    h->pn_terms_exact = L.pn_terms_exact;
    h->pp_terms = L.pp_terms;
    h->nn_terms = L.nn_terms;
    return h;
}

std::shared_ptr<DeviceIndex> upload_index(const Basis& b, const Tables& t, int device) {
    return upload_index(*pack_index(b, t), device);
}

int64_t DeviceIndex::device_bytes() const {
    int64_t s = 0;
    for (const DBuf* d : {&grp_block[0], &grp_block[1], &grp_fp0[0], &grp_fp0[1], &row_block, &block_row0, &nP, &nN,
                          &tdims, &en_off, &pp_rowptr, &pp_col, &nn_slice_base, &nn_slice_off, &nn_slice_len,
                          &nn_col, &pair_ptr, &pair_bi, &pair_dm, &pair_ptab, &pair_ntab, &ptab_rowptr_base,
                          &ptab_rowptr, &ptab_ent_base, &ptab_ent, &ntab_slice_base, &ntab_slice_off,
                          &ntab_slice_len, &ntab_ent, &nperm, &veff_off, &veff_cols2})
        s += (int64_t)d->cap;
    return s;
}

int64_t DeviceHam::coefficient_capacity() const {
    int64_t s = 0;
    for (int k = 0; k < 2; k++) s += (int64_t)(ep[k].cap + en[k].cap + pp_val[k].cap + nn_val[k].cap + veff[k].cap);
    return s;
}

void DeviceHam::release_coefficients() {
    for (int k = 0; k < 2; k++) { ep[k].release(); en[k].release(); pp_val[k].release(); nn_val[k].release(); veff[k].release(); }
}

void bind_ham(DeviceHam& h, std::shared_ptr<DeviceIndex> idx, const HostCoefficients& c) {
    if (c.pattern != idx->pattern)
        throw std::invalid_argument("interaction has a different pp/nn sparsity pattern than the shared tables");
    SM_HIP(hipSetDevice(idx->device));
    h.device = idx->device;
    h.mdim = idx->mdim;
    h.idx = std::move(idx);
    h.ep[0].upload(c.ep); h.en[0].upload(c.en); h.pp_val[0].upload(c.pp_val); h.nn_val[0].upload(c.nn_val);
    h.veff[0].upload(c.veff);
    h.ep[1].upload(c.ep32); h.en[1].upload(c.en32); h.pp_val[1].upload(c.pp_val32); h.nn_val[1].upload(c.nn_val32);
    h.veff[1].upload(c.veff32);
    for (int prec = 0; prec < 2; prec++) {
        HvTables& A = h.view[prec];
        A = h.idx->view[prec];
        A.ep = h.ep[prec].p; A.en = h.en[prec].p; A.pp_val = h.pp_val[prec].p;
        A.nn_val = h.nn_val[prec].p; A.veff = h.veff[prec].p;
    }
    h.pn_terms = h.idx->pn_terms;
    h.pn_terms_exact = h.idx->pn_terms_exact;
    h.pp_terms = h.idx->pp_terms;
    h.nn_terms = h.idx->nn_terms;
}

std::unique_ptr<DeviceHam> attach_ham(std::shared_ptr<DeviceIndex> idx, const Basis& b, const Tables& t) {
    auto c = pack_coefficients(b, t);
    auto h = std::make_unique<DeviceHam>();
    SM_HIP(hipSetDevice(idx->device));
    SM_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    bind_ham(*h, std::move(idx), *c);
    return h;
}

std::unique_ptr<DeviceHam> upload_ham(const Basis& b, const Tables& t, int device) {
    return attach_ham(upload_index(b, t, device), b, t);
}

void update_coefficients(DeviceHam& h, const Basis& b, const Tables& t) {
    auto c = pack_coefficients(b, t);
    bind_ham(h, h.idx, *c);
}

std::vector<std::pair<std::string, int64_t>> DeviceHam::table_bytes(int prec) const {
    const size_t s = prec == 0 ? 8 : 4;
    const DeviceIndex& I = *idx;
    const DBuf &grp_block = I.grp_block[prec], &grp_fp0 = I.grp_fp0[prec], &row_block = I.row_block,
               &block_row0 = I.block_row0, &nP = I.nP, &nN = I.nN, &tdims = I.tdims, &en_off = I.en_off,
               &pp_rowptr = I.pp_rowptr, &pp_col = I.pp_col, &nn_slice_base = I.nn_slice_base,
               &nn_slice_off = I.nn_slice_off, &nn_slice_len = I.nn_slice_len, &nn_col = I.nn_col,
               &pair_ptr = I.pair_ptr, &pair_bi = I.pair_bi, &pair_dm = I.pair_dm, &pair_ptab = I.pair_ptab,
               &pair_ntab = I.pair_ntab, &ptab_rowptr_base = I.ptab_rowptr_base, &ptab_rowptr = I.ptab_rowptr,
               &ptab_ent_base = I.ptab_ent_base, &ptab_ent = I.ptab_ent, &ntab_slice_base = I.ntab_slice_base,
               &ntab_slice_off = I.ntab_slice_off, &ntab_slice_len = I.ntab_slice_len, &ntab_ent = I.ntab_ent,
               &nperm = I.nperm,
               &veff_off = I.veff_off, &veff_cols2 = I.veff_cols2;
    std::vector<std::pair<std::string, int64_t>> v = {
        {"row_block", (int64_t)row_block.bytes}, {"block_row0", (int64_t)block_row0.bytes},
        {"nP", (int64_t)nP.bytes}, {"nN", (int64_t)nN.bytes}, {"tdims", (int64_t)tdims.bytes},
        {"en_off", (int64_t)en_off.bytes},
        {"ep", (int64_t)(ep[prec].bytes)}, {"en", (int64_t)(en[prec].bytes)},
        {"pp_rowptr", (int64_t)pp_rowptr.bytes}, {"pp_col", (int64_t)pp_col.bytes},
        {"pp_val", (int64_t)pp_val[prec].bytes},
        {"nn_slice_base", (int64_t)nn_slice_base.bytes}, {"nn_slice_off", (int64_t)nn_slice_off.bytes},
        {"nn_slice_len", (int64_t)nn_slice_len.bytes}, {"nn_col", (int64_t)nn_col.bytes},
        {"nn_val", (int64_t)nn_val[prec].bytes},
        {"pair_ptr", (int64_t)pair_ptr.bytes}, {"pair_bi", (int64_t)pair_bi.bytes},
        {"pair_dm", (int64_t)pair_dm.bytes}, {"pair_ptab", (int64_t)pair_ptab.bytes},
        {"pair_ntab", (int64_t)pair_ntab.bytes},
        {"ptab_rowptr_base", (int64_t)ptab_rowptr_base.bytes}, {"ptab_rowptr", (int64_t)ptab_rowptr.bytes},
        {"ptab_ent_base", (int64_t)ptab_ent_base.bytes}, {"ptab_ent", (int64_t)ptab_ent.bytes},
        {"ntab_slice_base", (int64_t)ntab_slice_base.bytes}, {"ntab_slice_off", (int64_t)ntab_slice_off.bytes},
        {"ntab_slice_len", (int64_t)ntab_slice_len.bytes}, {"ntab_ent", (int64_t)ntab_ent.bytes},
        {"nperm", (int64_t)nperm.bytes},
        {"veff_off", (int64_t)veff_off.bytes}, {"veff_cols2", (int64_t)veff_cols2.bytes},
        {"veff", (int64_t)veff[prec].bytes},
        {"vector_in", (int64_t)(mdim * s)}, {"vector_out", (int64_t)(mdim * s)}};
    return v;
}

}  // namespace smgpu
