/*
 * CPU oracle for NextLA.jl's recursive triangular solve/multiply
 * (unified_rectrxm!, src/rectrxm.jl:43-198; GEMM_ADD!/GEMM_SUB!, src/matmul.jl:5-80;
 * TRSM leaves, src/trsm.jl:5-150; TRMM leaves, src/trmm.jl:43-393).
 *
 * TEST INFRASTRUCTURE ONLY.  This restatement is the checker for the HIP
 * library; it is never linked into, loaded by or called from the product path.
 * Only tests/, the smoke check and the bench's CPU-baseline leg use it.
 *
 * Pinning (what makes this restatement trustworthy):
 *   - every (side, uplo, trans, func) combination is checked against OpenBLAS
 *     0.3.29 dtrsm/dtrmm/strsm/strmm (the library and version the reference's
 *     own tests compare against, test/unified_rectrxm.jl:36-40,
 *     test/Manifest.toml:494) through committed golden fixtures;
 *   - closed forms (identity, diagonal, unit bidiagonal, nilpotent series),
 *     side/uplo duality, linearity in alpha, TRMM->TRSM round trips;
 *   - exact rational arithmetic on tiny dyadic inputs for the split rule and
 *     the update ordering, including the non-power-of-two split branch.
 * Built with -ffp-contract=off so products and sums round separately, as in
 * Julia's default (non-fastmath) arithmetic.
 */
#include <stdlib.h>
#include <string.h>

#define OT double
#define OSF d
#include "rectrxm_oracle_impl.h"
#undef OT
#undef OSF

#define OT float
#define OSF s
#include "rectrxm_oracle_impl.h"
#undef OT
#undef OSF

int oracle_abi_version(void) { return 1; }
