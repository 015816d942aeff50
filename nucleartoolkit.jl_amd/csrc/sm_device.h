// Device-side objects shared by the HIP translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "sm_host.h"

namespace smgpu {

#define SM_HIP(call)                                                                  \
    do {                                                                              \
        hipError_t e_ = (call);                                                       \
        if (e_ != hipSuccess)                                                         \
            throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(e_) + \
                                     " at " + __FILE__ + ":" + std::to_string(__LINE__)); \
    } while (0)

// Device buffer owned by a handle.  `bytes` is the size in use, `cap` the size
// allocated; ensure() and upload() reuse the allocation when it is large enough.
struct DBuf {
    void* p = nullptr;
    size_t bytes = 0;
    size_t cap = 0;
    bool owned = true;         // false: a view of memory owned elsewhere (never freed here)
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf();
    void alloc(size_t n);      // fresh allocation of exactly n bytes
    void release();
    void ensure(size_t n) {    // grow-only reuse
        if (cap < n || !p) alloc(n);
        else bytes = n;
    }
    void view(void* ptr, size_t n) {   // point at n bytes of caller-owned device memory
        release();
        p = ptr;
        bytes = cap = n;
        owned = false;
    }
    template <typename T> void upload(const std::vector<T>& v) {
        ensure(v.size() * sizeof(T));
        if (!v.empty()) SM_HIP(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    }
    template <typename T> T* as() const { return (T*)p; }
};

// Kernel view of the Hamiltonian tables for one precision (all device pointers).
struct HvTables {
    int nb, R, G;
    const int32_t* row_block;   // R
    const int32_t* block_row0;  // nb
    const int32_t* grp_block;   // G: block of each row group
    const int32_t* grp_fp0;     // G: first proton row of the group
    const int32_t* nP;          // nb
    const int32_t* nN;          // nb
    const int64_t* tdims;       // nb+1
    const void* ep;             // R      (T)
    const int64_t* en_off;      // nb
    const void* en;             // sum nN (T)
    const int64_t* pp_rowptr;   // R+1
    const int32_t* pp_col;
    const void* pp_val;         // (T)
    const int64_t* nn_slice_base;  // nb: first slice of the block
    const int64_t* nn_slice_off;   // entry offset of a slice
    const int32_t* nn_slice_len;
    const int32_t* nn_col;
    const void* nn_val;            // (T)
    const int32_t* pair_ptr;       // nb+1
    const int32_t* pair_bi;
    const int32_t* pair_dm;
    const int32_t* pair_ptab;
    const int32_t* pair_ntab;
    const int64_t* ptab_rowptr_base;  // per p table: offset into ptab_rowptr
    const int32_t* ptab_rowptr;       // relative entry offsets
    const int64_t* ptab_ent_base;
    const uint32_t* ptab_ent;         // i | op<<20 | neg<<31
    const int64_t* ntab_slice_base;   // per n table: first slice
    const int64_t* ntab_slice_off;
    const int32_t* ntab_slice_len;
    const uint32_t* ntab_ent;         // in | col<<20 ; col = 2*bd + neg, pad col = 2*cols
    const int32_t* nperm;             // per block (offset en_off): walk slot -> neutron SD
    const int64_t* veff_off;          // per dM
    const int32_t* veff_cols2;        // per dM: 2*(cols+1)
    const void* veff;                 // (T) signed copies: [row][2*bd+neg]
    int rg;         // proton rows per workgroup
    int pch;        // hop slots per row and staging round (multiple of 4)
    int nN_max;
    int cols2_max;
    int npt;        // columns per lane: ceil(nN_max / tpb), rounded up to an instance
    int tpb;        // threads per workgroup (256, 512 or 1024)
    size_t lds_bytes;
};

// Interaction-independent device tables of one (model space, particle numbers,
// 2M): basis blocks, hop tables, row groups.  Shared by every interaction on
// that basis (interaction variants, serving requests of the same nucleus).
struct DeviceIndex {
    int device = 0;
    int64_t mdim = 0;
    int nb = 0, R = 0;
    DBuf grp_block[2], grp_fp0[2];   // row groups per precision (cost order)
    DBuf row_block, block_row0, nP, nN, tdims, en_off, pp_rowptr, pp_col, nn_slice_base, nn_slice_off,
        nn_slice_len, nn_col, pair_ptr, pair_bi, pair_dm, pair_ptab, pair_ntab, ptab_rowptr_base,
        ptab_rowptr, ptab_ent_base, ptab_ent, ntab_slice_base, ntab_slice_off, ntab_slice_len, ntab_ent, nperm,
        veff_off, veff_cols2;
    HvTables view[2];                // index part (coefficient pointers unset)
    uint64_t pattern = 0;            // fingerprint of the pp / nn sparsity patterns
    int64_t pn_terms = 0, pp_terms = 0, nn_terms = 0, pn_terms_exact = 0;
    int64_t lds_elems[2][5] = {};    // per precision: LDS elements of a one-vector launch (lds_model)
    int64_t device_bytes() const;    // HBM held by the index tables
};

// Device-resident Hamiltonian: shared index tables + this interaction's coefficients.
struct DeviceHam {
    int device = 0;
    hipStream_t stream = nullptr;
    int64_t mdim = 0;
    std::shared_ptr<DeviceIndex> idx;
    DBuf ep[2], en[2], pp_val[2], nn_val[2], veff[2];   // index 0: FP64, 1: FP32
    HvTables view[2];
    // byte accounting of every array the H·v kernel streams, per precision
    std::vector<std::pair<std::string, int64_t>> table_bytes(int prec) const;
    int64_t pn_terms = 0, pp_terms = 0, nn_terms = 0;   // FMAs per H·v (padded n-table entries included)
    int64_t pn_terms_exact = 0;
    std::shared_ptr<void> trl_ws[2];   // Lanczos workspace per precision, reused across solves
    int twoJmax = -1;                  // J^2 handles: largest 2J of the basis
    int64_t coefficient_capacity() const;   // HBM held by the coefficient buffers
    void release_coefficients();
    ~DeviceHam();
};

// Host copies of the kernel-layout arrays, built on host threads ahead of an
// upload (the serving loop prepares them while other solves use the device).
struct HostIndex;          // interaction-independent index arrays
struct HostCoefficients;   // one interaction's coefficients, FP64 and FP32
std::shared_ptr<const HostIndex> pack_index(const Basis& b, const Tables& t);
int64_t device_bytes(const HostIndex& x);            // HBM that upload_index allocates
std::shared_ptr<DeviceIndex> upload_index(const HostIndex& x, int device);
std::shared_ptr<const HostCoefficients> pack_coefficients(const Basis& b, const Tables& t);
int64_t device_bytes(const HostCoefficients& c);     // HBM of the coefficients, both precisions
int64_t coefficient_growth(const DeviceHam& h, const HostCoefficients& c);   // bind_ham's new allocation
uint64_t pattern_of(const HostCoefficients& c);
uint64_t content_hash(const HostCoefficients& c);                 // of the coefficient values
bool same_coefficients(const HostCoefficients& a, const HostCoefficients& b);   // identical Hamiltonians
uint64_t pattern_of(const HostIndex& x);
// Point h at idx with coefficients c, reusing h's coefficient buffers when large
// enough (no allocation in steady state).  Throws if the sparsity patterns differ.
void bind_ham(DeviceHam& h, std::shared_ptr<DeviceIndex> idx, const HostCoefficients& c);

// fingerprint of the interaction-dependent sparsity (pp / nn rows) on basis b
uint64_t coefficient_pattern(const Basis& b, const Tables& t);
std::shared_ptr<DeviceIndex> upload_index(const Basis& b, const Tables& t, int device);
// new handle on an existing index; throws if t's pp / nn pattern differs from the index's
std::unique_ptr<DeviceHam> attach_ham(std::shared_ptr<DeviceIndex> idx, const Basis& b, const Tables& t);
std::unique_ptr<DeviceHam> upload_ham(const Basis& b, const Tables& t, int device);
// w = H v for one vector; prec 0 = FP64 (double*), 1 = FP32 (float*)
void launch_hv(const DeviceHam& h, int prec, const void* v, void* w, hipStream_t s);
// input/output vectors of one H·v launch (up to four, each in the launch precision)
struct HvVecs {
    const void* v[4];
    void* w[4];
    // per-vector coefficients (interaction variants on one index; used by the PV
    // kernels): e_p, e_n, pp values, nn values and V_eff of vector b's Hamiltonian
    const void* ep[4];
    const void* en[4];
    const void* pp[4];
    const void* nn[4];
    const void* veff[4];
    // captured Lanczos steps: v[b] is a basis and the input is v[b] + (*vsel) * vld
    const int* vsel;
    int64_t vld;
};
// w[b] = H v[b] for nvec arbitrary vectors, in launches of up to four that share
// the table loads (each result bit-identical to launch_hv)
void launch_hv_set(const DeviceHam& h, int prec, int nvec, const void* const* v, void* const* w, hipStream_t s);
// w[b] = H_b v[b] for Hamiltonians that share one index (interaction variants on one
// basis): launches of up to four share the index loads and entry decodes, each vector
// with its own coefficients; results are bit-identical to launch_hv(*hams[b], ...)
void launch_hv_variants(const DeviceHam* const* hams, int prec, int nvec, const void* const* v, void* const* w,
                        hipStream_t s);
// w_b = H v_b for nvec vectors at v + b*ldv -> w + b*ldw (elements), in launches of
// up to four vectors that share the table loads
void launch_hv_multi(const DeviceHam& h, int prec, int nvec, const void* v, int64_t ldv, void* w, int64_t ldw,
                     hipStream_t s);
// one-vector H·v of column *vsel of the basis v (leading dimension ld) into w, the
// column read on the device (captured Lanczos steps)
void launch_hv_column(const DeviceHam& h, int prec, const void* v, int64_t ld, const int* vsel, void* w, hipStream_t s);
// re-upload interaction-dependent coefficients (same basis and index tables)
void update_coefficients(DeviceHam& h, const Basis& b, const Tables& t);

// Moment operators (M1, E2) of main_sm's calc_moment on the device: per operator
// and species the symmetric matrices of every M-set, flattened (rows of set s start
// at rp_base[s] in rowptr, their entries at ent_base[s]).
struct DeviceMoments {
    int device = 0;
    hipStream_t stream = nullptr;
    int64_t mdim = 0;
    int nb = 0, R = 0, twoM = 0;
    double bpar = 0.0;
    DBuf row_block, block_row0, nN, tdims, pset, nset;
    DBuf rp_base[2][2], ent_base[2][2], rowptr[2][2], col[2][2], val[2][2];   // [op][species]
    DBuf v, w, part;                       // work vectors of eval_moment (grow-only)
    ~DeviceMoments();
};
std::unique_ptr<DeviceMoments> upload_moments(const MomentOps& mo, const Basis& b, int device);
// w = O v, O = M1 (which 0) or E2 (1), FP64 device vectors in the M-scheme order
void moments_apply(const DeviceMoments& m, int which, const double* v, double* w, hipStream_t s);
// eval_moment (transit.jl:702-765): mu[n], Q[n] of nvec host vectors with J values
// totJs (NaN where the reference prints nothing: J < 0 or a vanishing 3j symbol)
void eval_moment(DeviceMoments& m, int nvec, const double* vecs, const double* totJs, double* mu, double* Q);

// transit_main's operators on the device: right basis -> left basis
struct DeviceTransitions {
    int device = 0;
    hipStream_t stream = nullptr;
    int64_t mdim_l = 0, mdim_r = 0;
    int R = 0, twoM_l = 0, twoM_r = 0, mu2 = 0;
    double bpar = 0.0;
    DBuf row_block, block_row0, nN_l, tdims_l, pset_l, nset_l, srcP, srcN, nN_r, tdims_r;
    DBuf rp_base[2][2], ent_base[2][2], rowptr[2][2], col[2][2], val[2][2];   // [op][species]
    DBuf vr, wl, left, part;               // work vectors of transit (grow-only)
    ~DeviceTransitions();
};
std::unique_ptr<DeviceTransitions> upload_transitions(const TransitionOps& to, const Basis& left, const Basis& right,
                                                      int device);
// w_left = O v_right, O = M1 (which 0) or E2 (1), FP64 device vectors
void transitions_apply(const DeviceTransitions& t, int which, const double* v, double* w, hipStream_t s);
// transit_main (transit.jl:561-700): the reduced matrix elements <l|O_mu|r> / 3j(Jl lam Jr; -Ml mu Mr) of
// nl left and nr right host vectors, [l * nr + r], for M1 (lam 1) and E2 (lam 2); NaN where the
// transition is not allowed (triangle rule, mu above lam) or the 3j symbol vanishes
void transit(DeviceTransitions& t, int jl2, int jr2, int nl, const double* wl, int nr, const double* wr, double* m1,
             double* e2);

}  // namespace smgpu
