/*
 * su3.h -- SU(3) matrix types with the memory image of openQCD-1.4's su3.h
 * (the reference includes it at include/smear.h:4-6, include/wloop.h:8-10).
 * In a build against openQCD, openQCD's own su3.h is used instead: the
 * layout (9 complex doubles, row major c11..c33, 144 bytes) is identical.
 */
#ifndef SU3_H
#define SU3_H

typedef struct
{
   double re, im;
} complex_dble;

typedef struct
{
   complex_dble c1, c2, c3;
} su3_vector_dble;

typedef struct
{
   complex_dble c11, c12, c13, c21, c22, c23, c31, c32, c33;
} su3_dble;

#endif
