/*
 * ref_precise_shim.h - forced-include instrumentation for the oracle build of
 * the reference fdtd.c (oracle/Makefile, target ref-precise).
 *
 * The reference prints monitor and full-monitor rows with "%e" (7 significant
 * digits, fdtd.c:3754, 3839), too coarse for a 1e-10 parity check.  This header
 * is force-included ahead of fdtd.c (gcc -include) and rewrites every bare
 * "%e" conversion of fprintf into "%.17e", which round-trips an IEEE double.
 * The reference source itself is compiled unmodified from where it lies.
 * TEST INFRASTRUCTURE ONLY.
 */
#ifndef ORC_REF_PRECISE_SHIM_H
#define ORC_REF_PRECISE_SHIM_H
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

static int orc_precise_fprintf(FILE *f, const char *fmt, ...) {
    char buf[8192];
    size_t o = 0;
    for (const char *p = fmt; *p && o + 8 < sizeof(buf); p++) {
        if (p[0] == '%' && p[1] == '%') {
            buf[o++] = *p++;
            buf[o++] = *p;
        } else if (p[0] == '%' && p[1] == 'e') {
            memcpy(buf + o, "%.17e", 5);
            o += 5;
            p++;
        } else {
            buf[o++] = *p;
        }
    }
    buf[o] = 0;
    va_list ap;
    va_start(ap, fmt);
    int r = vfprintf(f, buf, ap);
    va_end(ap);
    return r;
}
#define fprintf orc_precise_fprintf
#endif
