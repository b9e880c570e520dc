/* vpmfmm_reference_velocity_calc.hpp -- takes the place of the reference's
 * src/velocity_calculation/velocity_calc.hpp (copy this file over it), so
 * that every translation unit of the reference program that includes that
 * header -- main.cpp:44 (velocity_tool, main.cpp:72, 410) and
 * src/penalization/penalization_calc.cpp:3 (the per-call VelocityCalc of
 * penalization::no_slip_iterative, :363, :402) -- sees one VelocityCalc, the
 * GPU drop-in of vpmfmm_velocity_calc.hpp.  src/velocity_calculation/
 * velocity_calc.cpp, which defines the replaced class's members, leaves the
 * build.  Compile with -I<this directory>.
 *
 * It keeps the replaced header's include guard and the headers it brought in
 * (Utils.hpp, treeCell.hpp, gridNode.hpp, poissonLSMPS.hpp), so code that
 * relied on them through it still compiles.  The relative paths hold at the
 * replaced header's place in the reference tree. */
#ifndef INCLUDED_VELOCITY_POISSON
#define INCLUDED_VELOCITY_POISSON

#include "../../Utils.hpp"
#include "../FMM/treeCell.hpp"
#include "../grid_block/gridNode.hpp"
#include "../LSMPS_poisson/poissonLSMPS.hpp"

#include "vpmfmm_velocity_calc.hpp"

#endif
