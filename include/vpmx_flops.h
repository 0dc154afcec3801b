/*
 * vpmx_flops.h -- frozen algorithmic work counts used for roofline reports.
 *
 * Convention: add, sub, mul = 1 flop; FMA = 2; div and sqrt = 1.
 * Transcendentals (exp, erf, rsqrt-as-special) are counted separately and not
 * credited as flops.  Counts follow the reference pair-loop form
 * (src/FLOWUnsteady_processing_force.jl:895-922) extended to J = grad U:
 *   dX 3; r^2 5; sqrt + div 2; rho 1; kernel 8 (gaussian-erf) / 12
 *   (Winckelmans); dX x Gamma 9; -1/(4 pi r^3) 3; g a 1; U 6; radial aux 5+4;
 *   J radial 18; J Kronecker 12.
 * Precision overheads (relative coordinates, FP64 folds) are not credited.
 */
#ifndef VPMX_FLOPS_H
#define VPMX_FLOPS_H

#define VPMX_FLOPS_UJ_PAIR_GAUSSIANERF 77
#define VPMX_FLOPS_UJ_PAIR_WINCKELMANS 81
#define VPMX_FLOPS_ESTR_PAIR 42
/* Fused FP64 RK substep: read X, Gamma, sigma, U, J, SFS, C, M (30 doubles),
 * write X, Gamma, sigma, M (14 doubles). */
#define VPMX_BYTES_RK_SUBSTEP_PER_PARTICLE 352

/* MI355X published peaks used as roofline denominators. */
#define VPMX_PEAK_FP32_VECTOR_TFLOPS 157.3
#define VPMX_PEAK_FP64_VECTOR_TFLOPS 78.6
#define VPMX_PEAK_HBM_GBS 8000.0

#endif /* VPMX_FLOPS_H */
