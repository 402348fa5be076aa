// C entry point over the reference's fastgl.cpp (compiled where it lies, into
// oracle/_ref/; test oracle only).  Mirrors legendre/fastgl/fastgl_wrap.pyx:22-29
// (GLPairs): x[i] = GLPair(N, N-i).x(), w[i] = weight, i.e. ascending nodes.
#include "fastgl.h"

extern "C" void ref_leggauss(int N, double* x, double* w)
{
    for (int i = 0; i < N; ++i) {
        fastgl::QuadPair p = fastgl::GLPair((size_t)N, (size_t)(N - i));
        x[i] = p.x();
        w[i] = p.weight;
    }
}
