// The single-grid sweep kernels in the fast arithmetic (F = true): sweep_kernels.hip compiled a
// second time with FMA contraction allowed. The fast arithmetic is held to the 1e-10 accuracy
// bar rather than to the reference's bits, and a fused a * b + c (one rounding) takes one
// instruction where the exact kernels spend two; the batch's fast kernels (batch_fast.hip) are
// compiled the same way, so a grid in a fast batch and the same grid alone end on the same bits.
#pragma clang fp contract(fast)
#define CFD_FAST_TU 1
#include "sweep_kernels.hip"
