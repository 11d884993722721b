/* bf_internal.h -- helpers shared by the C host files of libdsaxbf. */
#ifndef BF_INTERNAL_H
#define BF_INTERNAL_H
#include "dsaxbf.h"

/* Record a printf-style message as this thread's last error; returns code. */
int dsaxbf_fail(int code, const char *fmt, ...);
/* Clear this thread's last error. */
void dsaxbf_clear_error(void);
/* Validate cfg; returns DSAXBF_OK or DSAXBF_EINVAL with a message. */
int dsaxbf_check_config(const dsaxbf_config *cfg);

/* Internals of a plan with weights, for the batch layer (bf_batch.c). */
struct bfk_dims;
int dsaxbf_plan_internals(dsaxbf_plan *p, struct bfk_dims *dims, void **d_wfrag, float **d_wscale,
                          float **d_bpart);
/* Time each variant the plan's shape supports on the given device buffers
 * (reps launches after one warm-up);
 * ms[v] = mean ms or -1, *best = fastest.  The plan's variant is unchanged. */
int dsaxbf_time_variants(dsaxbf_plan *p, const uint8_t *d_in, const float *d_scale, uint8_t *d_out,
                         int reps, float *ms, int *best);

#endif
