// CPU ORACLE — cone-traced lighting (test infrastructure only); filled in below.
#pragma once
#include "oracle.hpp"
