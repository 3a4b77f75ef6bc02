// devpow.cpp -- the device library's pow (hydrocode-fusion_amd/csrc/fastpow.hpp)
// compiled for the host, for liboracle_devpow.so.  TEST INFRASTRUCTURE ONLY.
#include "fastpow.hpp"

extern "C" double orc_devpow(double x, double y) { return hgrp::fast_pow(x, y); }
