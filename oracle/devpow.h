/* devpow.h -- forced into the oracle's sources for liboracle_devpow.so
 * (oracle/Makefile): pow() becomes the device library's pow, built for the
 * host.  TEST INFRASTRUCTURE ONLY. */
#include <math.h>
double orc_devpow(double x, double y);
#define pow(x, y) orc_devpow((x), (y))
