// variant.hpp -- which arithmetic variant this translation unit compiles.
//
// The kernel sources classic2.hip, classic_split.hip, ensemble.hip and sharp2.hip
// are compiled twice (pyclaw_amd/_build.py):
//   exact (CLAW_FAST 0): no floating-point contraction, so every FP64 expression
//     rounds as the reference's Fortran does (CLAWHIP_F64, bit for bit);
//   fast (CLAW_FAST 1): a*b + c written in one expression becomes one FMA
//     (`#pragma clang fp contract(on)`, before any header is parsed, so the
//     templates of riemann.hpp / unsplit.hpp / common.hpp are contracted too):
//     the FP32 + Kahan paths (CLAWHIP_F32K) and CLAWHIP_F64_FAST, which are held
//     to tolerances.  Contraction is decided per source expression, so every
//     kernel built from the same expressions rounds the same way (an ensemble
//     member equals the same problem stepped alone).
// Each variant's kernels and launchers live in namespace claw::exact or
// claw::fast; dispatch.hip routes an entry point by its dtype.  The step
// controllers (stepctl.hpp, sharpctl.hpp) switch contraction off in their bodies:
// their FP64 control arithmetic equals the host's in either variant.
// Include this first, before any other header.
#pragma once

#ifndef CLAW_FAST
#define CLAW_FAST 0
#endif

#if CLAW_FAST
#pragma clang fp contract(on)
#define CLAW_VNS fast
#else
#define CLAW_VNS exact
#endif
