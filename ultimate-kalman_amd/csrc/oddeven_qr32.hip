/*
 * oddeven_qr32.hip -- the odd-even QR phase kernels of bucket 32 (state
 * dimensions 25..32) in every launch variant, FP64 and FP32.
 */
#include "oddeven_qr.h"

UKQR_DEFINE(32)
