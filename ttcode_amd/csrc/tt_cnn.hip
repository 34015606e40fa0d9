// tt_cnn.hip -- learned initial guess for the FLD solve (new; no reference
// counterpart).  Entry points are declared in include/ttgpu.h; until weights
// are set the energy solve never uses a learned guess.
#include "tt_common.h"

namespace tt {

static bool g_cnn_ready = false;

void cnn_free() { g_cnn_ready = false; }
bool cnn_ready() { return g_cnn_ready; }
int cnn_predict_guess(double) { return fail(1, "learned guess: no weights set"); }

}  // namespace tt

using namespace tt;

extern "C" {
int64_t tt_cnn_num_weights(void) { return 0; }
int64_t tt_cnn_num_biases(void) { return 0; }
int tt_cnn_set_weights(const float*, int64_t, const float*, int64_t) { return fail(1, "not available"); }
int tt_cnn_get_weights(float*, int64_t, float*, int64_t) { return fail(1, "not available"); }
int tt_cnn_forward_host(int, int, int, const float*, float*) { return fail(1, "not available"); }
int tt_cnn_train_step_host(int, int, int, const float*, const float*, float, float*, float*, float*) {
    return fail(1, "not available");
}
}
