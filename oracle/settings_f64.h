/*
 * TEST INFRASTRUCTURE. Forced-include configuration for building the
 * reference library with its scalar type set to double (the FP64 oracle).
 *
 * The reference selects its precision in common/settings.h: `typedef float
 * scalar` plus three constants that the header itself says must follow the
 * scalar type (B3_MAX_SCALAR, B3_EPSILON at settings.h:37, 49, 54; B3_PI is
 * formed in the scalar type). This file defines that header's include guard
 * (settings.h:19-20) first, so the reference's own header is skipped, and
 * declares the same names with scalar = double. Every other constant keeps
 * the value settings.h gives it. Used only by oracle/Makefile's ref64 target.
 */
#ifndef B3_SETTINGS_H
#define B3_SETTINGS_H

#include <assert.h>
#include <cstring>
#include <new>
#include <float.h>

typedef signed char int8;
typedef signed short int16;
typedef signed int int32;
typedef signed long long int64;
typedef unsigned char uint8;
typedef unsigned short uint16;
typedef unsigned int uint32;
typedef unsigned long long uint64;

typedef double scalar;
typedef double scalar64;

#define B3_MAX_U8 (0xFF)
#define B3_MAX_U32 (0xFFFFFFFF)
#define B3_MAX_SCALAR (DBL_MAX)
#define B3_EPSILON (DBL_EPSILON)
#define B3_PI scalar(3.14159265358979323846)

#define B3_AABB_EXTENSION scalar(0.2)
#define B3_LINEAR_SLOP scalar(0.005)
#define B3_AABB_MULTIPLIER scalar(2)
#define B3_MAX_TRANSLATION scalar(2.0)
#define B3_MAX_TRANSLATION_SQUARED (B3_MAX_TRANSLATION * B3_MAX_TRANSLATION)
#define B3_CONTACT_STIFFNESS scalar(1000.0)
#define B3_CONTACT_DAMPING_STIFFNESS scalar(0.0)
#define B3_MAX_CONTACT_LINEAR_CORRECTION scalar(0.2)
#define B3_BAUMGARTE scalar(0.2)

#define B3_NOT_USED(x) ((void)(x))
#define B3_ASSERT(c) assert(c)
#define B3_STATIC_ASSERT(c) static_assert(c)

#define B3_KiB(n) (1024 * n)
#define B3_MiB(n) (1024 * B3_KiB(n))
#define B3_GiB(n) (1024 * B3_MiB(n))

#ifndef B3_FORCE_INLINE
#define B3_FORCE_INLINE __inline
#endif

void* b3Alloc(uint32 size);
void b3Free(void* block);
void b3Log(const char* string, ...);

struct b3Version {
	uint32 major;
	uint32 minor;
	uint32 revision;
};

extern b3Version b3_version;

#endif
