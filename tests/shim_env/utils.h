/* Stand-in for src/core/utils.h (timer_info and the accumulated GPU time).  Test only. */
#ifndef MXR_SHIM_ENV_UTILS_H
#define MXR_SHIM_ENV_UTILS_H
#include <sys/time.h>
typedef struct timer_info {
  struct timeval start;
  struct timeval stop;
} timer_info;
extern float timer_gpu;
void timer_start(timer_info* timer);
void timer_stop(timer_info* timer);
float timer_delta(timer_info timer);
#endif
