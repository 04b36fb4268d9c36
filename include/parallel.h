/*
 * parallel.h -- thread-control hints kept linkable for drop-in clients.
 *
 * The reference exposes a four-function parallel-primitive layer
 * (c/parallel.h:31-37) whose TBB and sequential backends run the smoothers'
 * per-step callbacks on host threads. In libukhip that layer is replaced by
 * the HIP kernels behind kalman_smooth_oddeven/associative; the two set-up
 * calls clients make (e.g. c/performance.c:321-322) are kept as hints:
 *
 *  parallel_set_thread_limit(n)  -- n > 0 limits the host threads the library
 *                                  uses for packing/unpacking step arrays.
 *  parallel_set_blocksize(b)     -- b > 0 is recorded and reported by
 *                                  ukhip_get_blocksize(); the GPU work split is
 *                                  chosen by the library's own tuner.
 *
 * foreach_in_range*, prefix_sums_pointers and the spin mutex take host
 * callbacks over host pointer arrays and are not part of this library.
 */
#ifndef UKHIP_PARALLEL_H
#define UKHIP_PARALLEL_H

#ifdef __cplusplus
extern "C" {
#endif

void parallel_set_thread_limit(int number_of_threads); /* c/parallel.h:31 */
void parallel_set_blocksize(int blocksize_in);          /* c/parallel.h:32 */

#ifdef __cplusplus
}
#endif

#endif /* UKHIP_PARALLEL_H */
