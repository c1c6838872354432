/* Drop-in replacement for the reference's include/definitions.h:3-5.
 * Same macro names and values, so C code written against the reference
 * compiles unchanged.  HOST_DEVICE is not needed by this library's headers. */
#pragma once

#define DIMENSIONS 3
#define STENCIL 4
#define NEIGHBOURHOOD 27
