/*
 * oracle/pdhg_enh.c -- TEST INFRASTRUCTURE ONLY.
 * CPU port of the engine's enhanced mode (filled in with the GPU enhanced path).
 */
#include "pdhg_oracle.h"
