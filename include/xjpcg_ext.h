/*
 * Extensions of the pcg.h handle API (libxjpcg.so).
 */
#ifndef XJPCG_EXT_H
#define XJPCG_EXT_H

#include <stdint.h>

#include "pcg.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Per-iteration history of the last solve on `handle`: 64-byte records laid
 * out as the reference's CGSolverInstr (pcg/hw/include/impl/cgInstr.hpp:117-132,
 * little-endian u32 ID, Iter, VecSize, MaxIter; f64 Tols, Res, RZ, Alpha,
 * Beta; u64 Clock).  Record 0 is the set-up record the reference host writes
 * (cgVector.hpp:149-163); record k is iteration k as krnl_control stores it
 * (control.hpp:75-90), with Clock from the GPU's 100 MHz wall clock.
 * The reference reads these records back but exposes only the final residual
 * and count (pcgImp.hpp:171-180).  Copies min(count, capacity) records to
 * `records` (may be NULL to query) and sets *count = executed iterations + 1;
 * *count = 0 when the last solve kept no history (maxIter + 2 > 2^20). */
XJPCG_Status_t xJPCG_getHistory(const XJPCG_Handle_t* handle, void* records, uint64_t capacity, uint64_t* count);

#ifdef __cplusplus
}
#endif
#endif
