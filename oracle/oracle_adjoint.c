/*
 * oracle_adjoint.c -- CPU restatement of the adjoint (reverse-time) sweep.
 *
 * TEST INFRASTRUCTURE ONLY (see pic_oracle.c).  Filled in with the adjoint
 * kernels; this translation unit is intentionally empty for now.
 */
int or_adjoint_version(void) { return 0; }
