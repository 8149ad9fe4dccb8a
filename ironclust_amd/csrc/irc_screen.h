// irc_screen.h -- MFMA distance screen for the kNN pass (gfx950).
#pragma once
#include "irc_common.h"

namespace irc {

struct ScreenBufs {
    void* pack = nullptr;     // packed candidate operands (fp32 or bf16 hi/lo rows)
    void* aux = nullptr;      // per-block norms / bounds
    int* fallback = nullptr;  // queries whose candidate set was not certified
    int* n_fallback = nullptr;
};

bool screen_supported(const DevPlan& p);
size_t screen_pack_bytes(const DevPlan& p, int64_t n_cols, int prec);
size_t screen_aux_bytes(const DevPlan& p);
hipError_t launch_screen_knn(hipStream_t st, const DevPlan& p, int nq, int prec, int k_over, const ScreenBufs& sb,
                             int n_cu, int* knn_col, double* knn_t, double* dk2);

}  // namespace irc
