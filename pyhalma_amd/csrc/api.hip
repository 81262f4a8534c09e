// libhalma: error state, version, device queries and tuning knobs.
#include <stdarg.h>
#include <stdio.h>

#include "halma_common.h"
#include "sym.h"
#include "tuning.h"

namespace halma {

static thread_local char g_err[1024] = "";

int set_error(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

int hip_fail(hipError_t e, const char* what, const char* file, int line) {
    set_error(static_cast<int>(e), "HIP error %d (%s) in %s at %s:%d", static_cast<int>(e),
              hipGetErrorString(e), what, file, line);
    return static_cast<int>(e);
}

int device_cu_count() {
    static thread_local int cached_dev = -1;
    static thread_local int cached_cus = 0;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (dev != cached_dev) {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
            hipSuccess)
            return 0;
        cached_dev = dev;
        cached_cus = cus;
    }
    return cached_cus;
}

Tuning g_tuning;

}  // namespace halma

extern "C" {

const char* halma_last_error(void) { return halma::g_err; }

int halma_version(void) { return 10000; }

int halma_device_cu_count(void) { return halma::device_cu_count(); }

int halma_tune_set(int key, int value) {
    using halma::g_tuning;
    switch (key) {
        case HALMA_TUNE_TESTS_PER_LANE:
            if (value != 1 && value != 2 && value != 4)
                return halma::set_error(HALMA_E_ARG, "tests per lane must be 1, 2 or 4");
            g_tuning.tests_per_lane = value;
            return 0;
        case HALMA_TUNE_SUBTILE:
            if (value != 4 && value != 8)
                return halma::set_error(HALMA_E_ARG, "subtile must be 4 or 8");
            g_tuning.subtile = value;
            return 0;
        case HALMA_TUNE_MIN_CHUNK:
            if (value < 64 || value % 64)
                return halma::set_error(HALMA_E_ARG, "min chunk must be a multiple of 64");
            g_tuning.min_chunk = value;
            return 0;
        case HALMA_TUNE_ITEMS_PER_WAVE:
            if (value < 1 || value > 64)
                return halma::set_error(HALMA_E_ARG, "items per wave must be in [1, 64]");
            g_tuning.items_per_wave = value;
            return 0;
        case HALMA_TUNE_WAVES_PER_CU:
            if (value < 4 || value > 32 || value % 4)
                return halma::set_error(HALMA_E_ARG,
                                        "waves per CU must be a multiple of 4 in [4, 32]");
            g_tuning.waves_per_cu = value;
            return 0;
        case HALMA_TUNE_SYM_MAX:
            if (value < 0 || value > halma::kSymMaxMembers)
                return halma::set_error(HALMA_E_ARG, "symmetric-kernel halo limit must be in "
                                                     "[0, %d]", halma::kSymMaxMembers);
            g_tuning.sym_max = value;
            return 0;
        default:
            return halma::set_error(HALMA_E_ARG, "unknown tuning key %d", key);
    }
}

int halma_tune_get(int key) {
    using halma::g_tuning;
    switch (key) {
        case HALMA_TUNE_TESTS_PER_LANE: return g_tuning.tests_per_lane;
        case HALMA_TUNE_SUBTILE: return g_tuning.subtile;
        case HALMA_TUNE_MIN_CHUNK: return g_tuning.min_chunk;
        case HALMA_TUNE_ITEMS_PER_WAVE: return g_tuning.items_per_wave;
        case HALMA_TUNE_WAVES_PER_CU: return g_tuning.waves_per_cu;
        case HALMA_TUNE_SYM_MAX: return g_tuning.sym_max;
        default: return halma::set_error(HALMA_E_ARG, "unknown tuning key %d", key);
    }
}

}  // extern "C"
