// C++ host mirror of the reference's numeric entry point (include/numeric.h:6):
// LUonDevice(Symbolic_Matrix&, ostream& out, ostream& err, bool PERTURB)
// factors A_sym.val in place on the GPU (FP32), writes "Device 0: ..." and
// "Total GPU time: X ms" to `out` and error text to `err`, and returns early
// without touching val on failure (src/numeric.cu:347-639).
#ifndef GLU_MIRROR_NUMERIC_H
#define GLU_MIRROR_NUMERIC_H

#include <iostream>

#include "symbolic.h"

void LUonDevice(Symbolic_Matrix &A_sym, std::ostream &out, std::ostream &err, bool PERTURB);

#endif
