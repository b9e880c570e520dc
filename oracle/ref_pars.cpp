// ref_pars.cpp -- FMM knobs for the reference build in oracle/_ref.
//
// TEST INFRASTRUCTURE ONLY.  The reference reads exactly four Pars values on
// its FMM path (global.cpp:303 expTree, :311 P_max, :312 par_count_max,
// :313 src_count_max; treeCell.hpp:179-180, treeCell.cpp:1086,
// fmm2D.cpp:199).  global.cpp defines them as constants for the whole
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
}  // namespace Pars

extern "C" void ref_pars_set(int p, int max_par, int max_src, double exp_tree) {
    Pars::P_max = p;
    Pars::par_count_max = max_par;
    Pars::src_count_max = max_src;
    Pars::expTree = exp_tree;
}
