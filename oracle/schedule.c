/*
 * oracle/schedule.c -- instantiates the CPU schedule restatement
 * (oracle/schedule_impl.h) for FP32 and FP64.  Test infrastructure only.
 */
#include <float.h>
#include <math.h>
#include <stddef.h>
#include <stdlib.h>

#define T float
#define SUF _f32
#define SFMIN FLT_MIN
#include "schedule_impl.h"
#undef T
#undef SUF
#undef SFMIN

#define T double
#define SUF _f64
#define SFMIN DBL_MIN
#include "schedule_impl.h"
#undef T
#undef SUF
#undef SFMIN
