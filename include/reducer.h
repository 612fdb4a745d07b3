/*
 * reducer.h -- source-compatible Reducer (reference reducer.h:3-12) on the
 * MI355X library: one wave64 launch with FP64 accumulation per call
 * (n4_reduce_sum/max/min), instead of the reference's multi-pass 32-lane
 * shuffle reduction.  d_odata is accepted for compatibility and unused.
 */
#ifndef N4_REDUCER_H
#define N4_REDUCER_H

class Reducer {
public:
    Reducer() {}
    ~Reducer() {}
    static float reduce_sum_wrapper(int n, float* d_idata, float* d_odata);
    static float reduce_max_wrapper(int n, float* d_idata, float* d_odata);
    static float reduce_min_wrapper(int n, float* d_idata, float* d_odata);
};

#endif
