/*
 * Index macros of the storage contract shared with the reference
 * (include/arrays/{param,fluid,temperature}.h there).  The layout must be
 * identical so that arrays round-trip through the reference's NPY restart
 * files: row-major, x (I) fastest, one halo row on each side in y.  On one
 * GPU jsize == jtot.  The macros expect local variables itot, jsize and the
 * array pointer of the same (lower-case) name.
 */
#ifndef SNS_ARRAYS_H
#define SNS_ARRAYS_H

/* grid (param): xf [1:itot+1], xc [0:itot+1], dxf [1:itot], dxc [1:itot+1],
 * yf [1:jtot+1], yc [1:jtot] */
#define XF(I)  (xf[(I)-1])
#define XC(I)  (xc[(I)])
#define DXF(I) (dxf[(I)-1])
#define DXC(I) (dxc[(I)-1])
#define YF(J)  (yf[(J)-1])
#define YC(J)  (yc[(J)-1])
#define XF_MEMSIZE  (sizeof(double)*(itot+1))
#define XC_MEMSIZE  (sizeof(double)*(itot+2))
#define DXF_MEMSIZE (sizeof(double)*(itot))
#define DXC_MEMSIZE (sizeof(double)*(itot+1))
#define YF_MEMSIZE  (sizeof(double)*(jtot+1))
#define YC_MEMSIZE  (sizeof(double)*(jtot))

/* fields with y halos: ux [1:itot+1] x [0:jsize+1]; uy, p, psi, temp [0:itot+1] x [0:jsize+1] */
#define UX(I, J)   (ux[(J)*(itot+1)+((I)-1)])
#define UY(I, J)   (uy[(J)*(itot+2)+(I)])
#define P(I, J)    (p[(J)*(itot+2)+(I)])
#define PSI(I, J)  (psi[(J)*(itot+2)+(I)])
#define TEMP(I, J) (temp[(J)*(itot+2)+(I)])
#define UX_MEMSIZE   (sizeof(double)*(itot+1)*(jsize+2))
#define UY_MEMSIZE   (sizeof(double)*(itot+2)*(jsize+2))
#define P_MEMSIZE    (sizeof(double)*(itot+2)*(jsize+2))
#define PSI_MEMSIZE  (sizeof(double)*(itot+2)*(jsize+2))
#define TEMP_MEMSIZE (sizeof(double)*(itot+2)*(jsize+2))

/* source terms without halos: srcux* and tempforcex [2:itot] x [1:jsize];
 * srcuy* and srctemp* [1:itot] x [1:jsize] */
#define SRCUXA_MEMSIZE     (sizeof(double)*(itot-1)*(jsize))
#define SRCUYA_MEMSIZE     (sizeof(double)*(itot  )*(jsize))
#define TEMPFORCEX_MEMSIZE (sizeof(double)*(itot-1)*(jsize))
#define SRCTEMPA_MEMSIZE   (sizeof(double)*(itot  )*(jsize))

#endif
