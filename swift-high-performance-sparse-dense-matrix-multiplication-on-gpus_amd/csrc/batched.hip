// batched.hip -- many Swift SpMMs in one persistent launch (placeholder until
// the persistent kernel lands; every entry point reports UNSUPPORTED).
#include "swift_internal.h"

struct swift_batch {
    int count = 0;
};

extern "C" int swift_batch_create(swift_batch **out, swift_matrix *const *, const int *, int,
                                  swift_stream_t) {
    if (out) *out = nullptr;
    swift::set_error("swift_batch_create: not built yet");
    return SWIFT_ERR_UNSUPPORTED;
}
extern "C" int swift_batch_spmm(swift_batch *, const float *const *, float *const *,
                                swift_stream_t) {
    swift::set_error("swift_batch_spmm: not built yet");
    return SWIFT_ERR_UNSUPPORTED;
}
extern "C" int swift_batch_graph_launch(swift_batch *, const float *const *, float *const *,
                                        swift_stream_t) {
    swift::set_error("swift_batch_graph_launch: not built yet");
    return SWIFT_ERR_UNSUPPORTED;
}
extern "C" int swift_batch_params(const swift_batch *, int *, int *, int *) {
    return SWIFT_ERR_UNSUPPORTED;
}
extern "C" void swift_batch_destroy(swift_batch *b) { delete b; }
