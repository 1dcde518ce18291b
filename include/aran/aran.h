/*
 * aran.h -- library initialisation (reference: aran/aran.h:27-29).
 */
#ifndef ARAN_ARAN_H
#define ARAN_ARAN_H

#include <aran/arantypes.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void (*AranZeroFunc) (gpointer devel);

/* Binds the current HIP device (aran/aran.c:60-69).  Safe to call twice. */
void aran_init (void);

/* Message of the last failed call on this thread ("" when none). */
const gchar *aran_last_error (void);

#ifdef __cplusplus
}
#endif

#endif /* ARAN_ARAN_H */
