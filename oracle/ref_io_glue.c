/*
 * ref_io_glue.c -- caller-side glue for the reference's 2D I/O modules built
 * into oracle/_ref/libhydroref_io.so (test infrastructure only).  Owns config[]
 * as src/hydrocode_2D/hydrocode.c:123 does; the I/O functions themselves
 * (initialize_2D, file_2D_write, ...) are the reference's, exported unchanged.
 */
#include "var_struc.h"

double config[N_CONF];
