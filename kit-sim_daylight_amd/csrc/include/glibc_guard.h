/* Included first in every translation unit of the library: glibc's <time.h> declares a global `extern int daylight`
 * (XSI), which collides with the public namespace `daylight` of the C++ API. Declaring <time.h> once under another name
 * keeps both. */
#pragma once
#define daylight glibc_daylight_variable
#include <time.h>
#undef daylight
