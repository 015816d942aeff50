// Host-side data model of the M-scheme shell-model eigensolver.
//
// Everything here is built once on the host and uploaded to the device:
//   * Snt       — the parsed interaction file (readsmsnt semantics,
//                 src/ShellModel/shellmodel_main.jl:293-365)
//   * Species   — m-states, all Slater determinants (SDs) of one species grouped
//                 by 2M ("M-sets"), each set in the reference's order: descending
//                 bit integers, first m-state = most significant bit
//                 (shellmodel_main.jl:373-418, 732-848)
//   * Basis     — blocks (Mp ascending, Mn = 2M_tot - Mp), tdims
//   * Tables    — the device layout of H:
//       - one-body energy per SD (e_p, e_n),
//       - like-particle (pp / nn) two-body matrix per M-set as full symmetric
//         rows (merged duplicates) — the reference stores the upper triangle
//         and applies it twice (prep_2bjumps :1115-1137, operate_H! pp/nn),
//       - one-body hop tables per species and ordered M-set pair, gather form:
//         for every output SD f the inputs i with f = a+ c i, the operator
//         index of (a,c) within its dM list and the sign,
//       - V_eff[dM][ac][bd]: the pn coefficient of proton operator a+c times
//         neutron operator b+d, gathered in the direction the reference's pn1 /
//         pn2 loops apply it (lanczos_methods.jl:153-208).
// No part of this header is device code.
#pragma once
#include <array>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

namespace smgpu {

struct Orbit { int n, l, j2, tz; };

struct Snt {
    int lp = 0, ln = 0, cp = 0, cn = 0, massop = 0, A = 0;
    double Aref = 0.0, pw = 0.0;
    std::vector<Orbit> p_sps, n_sps;
    std::vector<double> spe_p, spe_n;
    // labels[ch] rows {a,b,c,d,J,line}; ch 0 = pp, 1 = nn, 2 = pn
    std::vector<std::array<int, 6>> labels[3];
    std::vector<double> tbmes[3];
};

// Throws std::runtime_error with a message on malformed input.
Snt read_snt(const std::string& path, const std::string& nucleus, bool ignore_massop);
// Element symbol -> Z (getZNA, shellmodel_main.jl:908-916).
int element_z(const std::string& symbol);

// "Cr48" + interaction core -> Z, N, valence numbers and 2M (twoM < 0: A % 2,
// the main_sm default, shellmodel_main.jl:110).  Throws std::invalid_argument.
struct NucleusSpec { int Z, N, vp, vn, twoM; };
NucleusSpec resolve_nucleus(const Snt& snt, const std::string& nucleus, int twoM);
// <j1/2 m1/2; j2/2 m2/2 | J/2 M/2>, correctly rounded (quad-precision Racah formula).
double clebsch_gordan(int j1, int m1, int j2, int m2, int J, int M);

struct Species {
    int L = 0;                       // m-states
    int nv = 0;                      // valence particles
    std::vector<int> orb, m2, j2;    // per m-state: orbit (0-based), 2m, 2j
    std::vector<int> Ms;             // M values of the M-sets, ascending
    std::vector<std::vector<uint64_t>> sets;  // SDs per M-set, descending
    std::vector<std::vector<double>> jocc;    // per set: per SD per orbit occupation (flattened)
    int norb = 0;
    int set_of_M(int M) const;
    int64_t find(int set, uint64_t sd) const;  // -1 if absent
};

struct Basis {
    Species sp[2];
    int twoM = 0;
    int nb = 0;
    std::vector<int> Mp, Mn;          // per block
    std::vector<int> pset, nset;      // M-set index per block and species
    std::vector<int64_t> tdims;       // nb + 1
    int64_t mdim = 0;
};

void build_species(Species& s, const std::vector<Orbit>& sps, int nv);
Basis build_basis(const Snt& snt, int vp, int vn, int twoM);

// One-body hop table, gather form, for output M-set `fset` from input M-set `iset`.
struct HopTable {
    int species = 0, iset = 0, fset = 0, dM = 0;   // dM = M_f - M_i
    std::vector<int64_t> rowptr;                   // |S_f| + 1
    std::vector<uint32_t> ent;                     // i | op<<20 | neg<<31
};

struct PairSym {            // symmetric like-particle matrix of one M-set, CSR
    std::vector<int64_t> rowptr;
    std::vector<int32_t> col;
    std::vector<double> val;
};

struct OneBodyOps {          // one-body operators a+ c of one species grouped by dM
    std::vector<int> dMs;                                  // ascending
    std::vector<std::vector<std::pair<int, int>>> ops;     // (a, c) m-state indices
    int index_of_dM(int dM) const;
};

struct Tables {
    std::vector<std::vector<double>> ep, en;      // per M-set per SD
    std::vector<PairSym> pp, nn;                  // per M-set
    OneBodyOps ob[2];
    std::vector<HopTable> hops;                   // proton and neutron tables
    std::map<std::array<int, 3>, int> hop_index;  // (species, iset, fset) -> hops index
    // V_eff per proton dM (index into ob[0].dMs): rows = proton ops, cols = neutron ops of -dM
    std::vector<std::vector<double>> veff;
    std::vector<int> veff_rows, veff_cols;
    // counts for reporting
    int64_t n_pp_ops = 0, n_nn_ops = 0, n_pn_ops = 0;
};

// The total angular momentum squared J^2 written as an interaction on the orbits
// of `model`: one-body j(j+1) per orbit and, for every two-particle state |ab;J>
// (pp, nn, pn), the diagonal matrix element <ab;J| 2 j1.j2 |ab;J> =
// J(J+1) - ja(ja+1) - jb(jb+1), since J^2 = sum_i j_i^2 + 2 sum_{i<j} j_i.j_j.
// Its tables on a basis give <v|J^2|v> with the H.v machinery (the quantity
// main_sm gets from prep_J / operate_J!, shellmodel_main.jl:149-157, 226-230).
Snt j2_operator(const Snt& model);

// ---- moments (transit.jl): the M = 0 components of the M1 and E2 one-body
// operators (prep_obtr_diag, :82-245) as symmetric matrices over the SDs of each
// M-set of a species (they conserve M), full rows with merged entries; op_obtr_diag
// (:247-283) applies each pushed term both ways, diagonal terms at half weight.
// gfactors = (g_l proton, g_l neutron, g_s proton, g_s neutron), effcharge = (e_p, e_n).
struct MomentOps {
    double bpar = 0.0;
    std::vector<PairSym> op[2][2];    // [0 = M1, 1 = E2][species] per M-set
};
MomentOps build_moment_ops(const Snt& snt, const Basis& b, double bpar, const double* gfactors,
                           const double* effcharge);
// The M1 and E2 one-body operators of transit_main (prep_obtr_nondiag, transit.jl:286-445):
// component mu = (2M_left - 2M_right) / 2 of the tensors, from the right basis into the left
// one (same interaction and nucleus).  op[mode][species][t] holds, for left M-set t, the CSR
// over its SDs of the source M-set at M_t - mu2 (not symmetric, no halving).
struct TransitionOps {
    double bpar = 0.0;
    int mu2 = 0;                      // 2M_left - 2M_right = m_a - m_c of every a+ c (units of 1/2)
    std::vector<PairSym> op[2][2];    // [0 = M1, 1 = E2][species] per left M-set
};
TransitionOps build_transition_ops(const Snt& snt, const Basis& left, const Basis& right, double bpar,
                                   const double* gfactors, const double* effcharge);
// oscillator length of main_sm (init_ho_by_mass, transit.jl:49-60, with formula 2
// for 16 <= A <= 40, shellmodel_main.jl:104-107)
double ho_bpar_main_sm(int A);
// (j1 j2 j3; m1 m2 m3) and {j1 j2 j3; j4 j5 j6}, doubled arguments
double wigner3j(int j1, int j2, int j3, int m1, int m2, int m3);
double wigner6j(int j1, int j2, int j3, int j4, int j5, int j6);
// <n1 l1|r^k|n2 l2> of oscillator orbits with length bpar (radius_power, transit.jl:62-80)
double radius_power(int k, int n1, int l1, int n2, int l2, double bpar);

// Interaction-dependent parts only (V_eff, pp, nn, e_p, e_n) can be rebuilt
// for a new interaction on the same basis with `build_coefficients`.
Tables build_tables(const Snt& snt, const Basis& b, int nthreads);
void build_coefficients(const Snt& snt, const Basis& b, Tables& t, int nthreads);

}  // namespace smgpu
