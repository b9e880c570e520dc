// ref_pars.cpp -- FMM knobs for the reference build in oracle/_ref.
//
// TEST INFRASTRUCTURE ONLY.  The reference reads four Pars values on its
// type-2 FMM path (global.cpp:303 expTree, :311 P_max, :312 par_count_max,
// :313 src_count_max; treeCell.hpp:179-180, treeCell.cpp:1086,
// fmm2D.cpp:199), and the "old FMM" block of global.cpp:316-328 on its
// type-1 path.  global.cpp defines them as constants for the whole
// simulation; this file defines the same four symbols as plain variables so a
// harness can sweep P and the leaf thresholds without touching the
// reference's files.  It deliberately does not include global.hpp (whose
// declarations carry `const`); namespace-scope variable names do not encode
// the type, so the reference objects bind to these definitions.
namespace Pars {
int P_max = 10;
int par_count_max = 75;
int src_count_max = 50;
double expTree = 1.0;
// The type-1 FMM (biotsavart_fmm_2d.cpp, subroutine_all_fmm_2d.cpp,
// regularization_function_2d.cpp) reads these (global.cpp:316-328).  npcm,
// nbmax, nbmrl and nbl1..nbl4 are array capacities there; they are set from
// lmax and ndp so no array overflows (the reference's fixed nbl1 = 8 lmax
// and nbl3 = 256 are smaller than the lists of strongly graded trees).
int lmax = 7;
int npcm = 100;
int nfwrd = 16;
int nbmax = 4 * (16384 - 1) / 3;
int nbmrl = 4 * (16384 - 1) / 3 + 5;
int nbl1 = 56;
int nbl2 = 256;
int nbl3 = 256;
int nbl4 = 16;
double tol2 = 1.0e-15;
int n_s = 50;
int ndp = 10;
int icutoff = 2;
}  // namespace Pars

extern "C" void ref_pars_set(int p, int max_par, int max_src, double exp_tree) {
    Pars::P_max = p;
    Pars::par_count_max = max_par;
    Pars::src_count_max = max_src;
    Pars::expTree = exp_tree;
}

extern "C" void ref_pars_set_type1(int n_s, int lmax, int ndp, int icutoff) {
    long long p4 = 1;
    for (int i = 0; i < lmax; ++i) p4 *= 4;
    Pars::lmax = lmax;
    Pars::npcm = ndp + 1;
    Pars::nbmax = (int)(4 * (p4 - 1) / 3);
    Pars::nbmrl = Pars::nbmax + 5;
    Pars::nbl1 = Pars::nbl2 = Pars::nbl3 = Pars::nbl4 = Pars::nbmax + 5;
    Pars::n_s = n_s;
    Pars::ndp = ndp;
    Pars::icutoff = icutoff;
}
