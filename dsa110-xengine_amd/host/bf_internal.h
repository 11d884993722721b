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

#endif
