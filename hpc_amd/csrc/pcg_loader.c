/*
 * dlopen loader for the xJPCG C API (built as libxjpcg_loader.a).
 *
 * Same role and error behaviour as the reference's pcg/sw/src/pcg_loader.c:
 * an application links this archive instead of the shared library, and each
 * pcg.h entry point resolves its implementation in libxjpcg.so on first use.
 * A failure to open the library or find a symbol returns
 * XJPCG_STATUS_DYNAMIC_LOADING_ERROR and stores a message that
 * xJPCG_getLastMessage returns, even for a NULL handle (pcg.h:228-233).
 * The single global message buffer is not thread-safe, as in the reference.
 *
 * The library is looked up as $XJPCG_LIBRARY when that is set, otherwise as
 * "libxjpcg.so" on the dynamic linker's search path.
 */
#include <dlfcn.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "pcg.h"

#define XJPCG_LOADER_MSG_LEN 4096

static char g_loader_msg[XJPCG_LOADER_MSG_LEN];
static int g_loader_failed = 0;

static const char* library_name(void) {
    const char* env = getenv("XJPCG_LIBRARY");
    return (env && *env) ? env : "libxjpcg.so";
}

static void loader_fail(const char* msg) {
    strncpy(g_loader_msg, msg, XJPCG_LOADER_MSG_LEN - 1);
    g_loader_msg[XJPCG_LOADER_MSG_LEN - 1] = 0;
    g_loader_failed = 1;
}

/* resolve `name` in the library, opening it if needed (RTLD_NOLOAD first, as the reference) */
static void* resolve(const char* name) {
    char msg[XJPCG_LOADER_MSG_LEN];
    const char* lib = library_name();
    void* so = dlopen(lib, RTLD_LAZY | RTLD_NOLOAD);
    if (!so) so = dlopen(lib, RTLD_LAZY | RTLD_GLOBAL);
    if (!so) {
        snprintf(msg, sizeof msg,
                 "ERROR: Cannot open library %s: %s.  Please ensure that the library's path is in LD_LIBRARY_PATH.\n",
                 lib, dlerror());
        loader_fail(msg);
        return NULL;
    }
    dlerror();
    void* fn = dlsym(so, name);
    const char* err = dlerror();
    if (err) {
        snprintf(msg, sizeof msg,
                 "ERROR: Cannot load symbol '%s': %s.  Possibly an older version of library %s is in use.\n", name,
                 err, lib);
        loader_fail(msg);
        return NULL;
    }
    g_loader_failed = 0;
    return fn;
}

XJPCG_Status_t xJPCG_createHandle(XJPCG_Handle_t** handle, const char* xclbinPath) {
    XJPCG_Status_t (*f)(XJPCG_Handle_t**, const char*);
    *(void**)&f = resolve("xJPCG_createHandle");
    return f ? f(handle, xclbinPath) : XJPCG_STATUS_DYNAMIC_LOADING_ERROR;
}

XJPCG_Status_t xJPCG_destroyHandle(XJPCG_Handle_t* handle) {
    XJPCG_Status_t (*f)(XJPCG_Handle_t*);
    *(void**)&f = resolve("xJPCG_destroyHandle");
    return f ? f(handle) : XJPCG_STATUS_DYNAMIC_LOADING_ERROR;
}

XJPCG_Status_t xJPCG_getMetrics(const XJPCG_Handle_t* handle, XJPCG_Metric_t* metric) {
    XJPCG_Status_t (*f)(const XJPCG_Handle_t*, XJPCG_Metric_t*);
    *(void**)&f = resolve("xJPCG_getMetrics");
    return f ? f(handle, metric) : XJPCG_STATUS_DYNAMIC_LOADING_ERROR;
}

XJPCG_Status_t xJPCG_peekAtLastStatus(const XJPCG_Handle_t* handle) {
    if (g_loader_failed) return XJPCG_STATUS_DYNAMIC_LOADING_ERROR;
    XJPCG_Status_t (*f)(const XJPCG_Handle_t*);
    *(void**)&f = resolve("xJPCG_peekAtLastStatus");
    return f ? f(handle) : XJPCG_STATUS_DYNAMIC_LOADING_ERROR;
}

const char* xJPCG_getLastMessage(const XJPCG_Handle_t* handle) {
    if (g_loader_failed) return g_loader_msg;
    const char* (*f)(const XJPCG_Handle_t*);
    *(void**)&f = resolve("xJPCG_getLastMessage");
    if (f) return f(handle);
    return g_loader_msg;
}

const char* xJPCG_getErrorString(const XJPCG_Status_t code) {
    if (code == XJPCG_STATUS_DYNAMIC_LOADING_ERROR) return "Dynamic loading error";
    const char* (*f)(XJPCG_Status_t);
    *(void**)&f = resolve("xJPCG_getErrorString");
    return f ? f(code) : "ERROR: Unable to get error string due to dynamic loading error.";
}

XJPCG_Status_t xJPCG_cscSymSolver(XJPCG_Handle_t* handle, const int64_t p_n, const int64_t p_nnz,
                                  const int64_t* p_rowIdx, const int64_t* p_colptr, const double* p_data,
                                  const double* p_diagA, const double* p_b, const double* p_x,
                                  const int64_t p_maxIter, const double p_tol, int64_t* p_iter, double* p_res,
                                  const XJPCG_Mode_t mode) {
    XJPCG_Status_t (*f)(XJPCG_Handle_t*, int64_t, int64_t, const int64_t*, const int64_t*, const double*,
                        const double*, const double*, const double*, int64_t, double, int64_t*, double*,
                        XJPCG_Mode_t);
    *(void**)&f = resolve("xJPCG_cscSymSolver");
    if (!f) return XJPCG_STATUS_DYNAMIC_LOADING_ERROR;
    return f(handle, p_n, p_nnz, p_rowIdx, p_colptr, p_data, p_diagA, p_b, p_x, p_maxIter, p_tol, p_iter, p_res,
             mode);
}

XJPCG_Status_t xJPCG_cooSolver(XJPCG_Handle_t* handle, const uint32_t p_n, const uint32_t p_nnz,
                               const uint32_t* p_rowIdx, const uint32_t* p_colIdx, const double* p_data,
                               const double* p_diagA, const double* p_b, const double* p_x, const uint32_t p_maxIter,
                               const double p_tol, uint32_t* p_iter, double* p_res, const XJPCG_Mode_t mode) {
    XJPCG_Status_t (*f)(XJPCG_Handle_t*, uint32_t, uint32_t, const uint32_t*, const uint32_t*, const double*,
                        const double*, const double*, const double*, uint32_t, double, uint32_t*, double*,
                        XJPCG_Mode_t);
    *(void**)&f = resolve("xJPCG_cooSolver");
    if (!f) return XJPCG_STATUS_DYNAMIC_LOADING_ERROR;
    return f(handle, p_n, p_nnz, p_rowIdx, p_colIdx, p_data, p_diagA, p_b, p_x, p_maxIter, p_tol, p_iter, p_res,
             mode);
}
