/*
 * arantypes.h -- the GLib scalar names used by the LibAran API, for builds
 * without GLib.  When glib.h is included first its definitions are used.
 */
#ifndef ARAN_TYPES_H
#define ARAN_TYPES_H

#include <stddef.h>
#include <stdint.h>

#ifndef __G_TYPES_H__
typedef void *gpointer;
typedef const void *gconstpointer;
typedef char gchar;
typedef int gint;
typedef unsigned int guint;
typedef long glong;
typedef unsigned long gulong;
typedef int gboolean;
typedef double gdouble;
typedef float gfloat;
typedef int8_t gint8;
typedef uint8_t guint8;
typedef int64_t gint64;
typedef uint64_t guint64;
typedef size_t gsize;
typedef ptrdiff_t gssize;
typedef void (*GFunc) (gpointer data, gpointer user_data);
#endif

#ifndef __G_TYPE_H__
typedef gsize GType;
#endif

#ifndef __G_NODE_H__
typedef enum
{
  G_IN_ORDER,
  G_PRE_ORDER,
  G_POST_ORDER,
  G_LEVEL_ORDER
} GTraverseType;
#endif

/* complex FP64 as LibAran's gcomplex128 (C99 double complex) */
typedef double _Complex gcomplex128;

#ifndef FALSE
#define FALSE (0)
#endif
#ifndef TRUE
#define TRUE (!FALSE)
#endif

#endif /* ARAN_TYPES_H */
