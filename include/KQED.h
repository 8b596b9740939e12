/*
 * KQED.h -- source-compatibility header for callers written against the reference library's main
 * header (reference src/HEADERS/KQED.h): the definitions those callers use (KQED_ALIGN, M_PI,
 * AlfQED) and the API of kqed.h, which covers the reference's kernel, kernel-combination and
 * lepton-loop headers.  The reference's timer helpers (KQED.h:155-163) are not part of this library.
 */
#ifndef KQED_COMPAT_H
#define KQED_COMPAT_H

#include <math.h>
#include <stdbool.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define KQED_ALIGNMENT (32)
#define KQED_ALIGN __attribute__((aligned(KQED_ALIGNMENT)))

#ifndef M_PI
#define M_PI (3.14159265358979323846)
#endif

/* QED coupling (reference KQED.h:45) */
static const double AlfQED = 1.0 / 137.035999;

#include "kqed.h"

#endif
