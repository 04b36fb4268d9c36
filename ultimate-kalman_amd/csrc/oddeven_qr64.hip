/*
 * oddeven_qr64.hip -- the odd-even QR phase kernels of bucket 64 (state
 * dimensions 48..64) in every launch variant, FP64 and FP32.
 */
#include "oddeven_qr.h"

UKQR_DEFINE(64)
