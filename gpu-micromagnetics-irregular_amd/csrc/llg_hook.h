// llg_hook.h -- library-internal: lets the ensemble layer take over a
// context's f() calls (LLG_Rhs, LLG_RhsResidual) so that the f of several
// instances can be evaluated by one LLG_RhsBatch.  Not part of the C ABI.
#pragma once

#include "llg/llg_batch.h"

// Called instead of launching f: y and ydot are the device arrays of the
// call, res the Newton-residual form or NULL.  Returns what f returns.
typedef int (*llg_rhs_hook)(void* data, llg_ctx* ctx, double* y, double* ydot,
                            const llg_res_in* res);

// fn = NULL restores the plain f.  Returns -1 for a NULL context.
extern "C" int LLG_SetRhsHook(llg_ctx* ctx, llg_rhs_hook fn, void* data);
