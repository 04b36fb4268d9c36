/*
 * oddeven_qr24.hip -- the odd-even QR phase kernels of bucket 24 (state
 * dimensions 17..24) in every launch variant, FP64 and FP32.
 */
#include "oddeven_qr.h"

UKQR_DEFINE(24)
