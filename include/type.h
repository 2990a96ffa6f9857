// Precision of the factors, as in the reference (include/type.h:3).
#ifndef GLU_MIRROR_TYPE_H
#define GLU_MIRROR_TYPE_H
typedef float REAL;
#endif
