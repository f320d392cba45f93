/*
 * orc_amr.c -- oracle for ghost filling, flux registers, reflux, flagging
 * (TEST INFRASTRUCTURE ONLY).
 */
#include "orc.h"
