/* SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0 */
/* Oracle build helper: glibc's <time.h> declares `extern int daylight`, which collides with the reference's namespace
 * `daylight`. Including <time.h> once under another name first keeps both. */
#define daylight glibc_daylight_variable
#include <time.h>
#undef daylight
