/*
 * oddeven_qr48.hip -- the odd-even QR phase kernels of bucket 48 (state
 * dimensions 32..48) in every launch variant, FP64 and FP32.
 */
#include "oddeven_qr.h"

UKQR_DEFINE(48)
