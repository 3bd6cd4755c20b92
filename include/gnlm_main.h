/*
 * gnlm_main.h -- C++ drop-in for the reference's host API.
 *
 * Replaces 3DNLM/main.h:4-8 and "2D NLM/11x11/main.h:4-7": one class
 * NLMeansProcessor whose static NL_Means overloads have the reference's exact
 * signatures, so the reference's drivers (3DNLM/nlm.cpp:65,
 * "2D NLM/11x11/nlm.cpp:36") compile unchanged against it and link with
 * -lgnlm instead of the CUDA objects.
 *
 *   3D: NL_Means(unsigned char* GPU_input, float* GPU_result)
 *       Geometry comes from the caller's config.h, as in the reference
 *       (3DNLM/config.h:4-9): SS (search window size), NS (patch size),
 *       IMAGE_SIZE_X, IMAGE_SIZE_Y, depth.  Include config.h before this
 *       header (or let this header include it: it does when __has_include
 *       finds one).  GPU_input: padded uint8 volume of
 *       (X+SS+NS-2)*(Y+SS+NS-2)*(depth+SS+NS-2) bytes, x fastest;
 *       GPU_result: X*Y*depth floats, x fastest.
 *   2D: NL_Means(unsigned char*, float*, int W, int H, int sr_size,
 *       int nb_size, int IMAGE_SIZE_X, int IMAGE_SIZE_Y) -- runtime geometry.
 *
 * Both are synchronous and host-pointer based, like the reference
 * (3DNLM/main.cu:7607-7680).  The filtering parameter is the reference's
 * PARAMETER = 10 on the 0-255 scale (3DNLM/main.cu:13); define
 * GNLM_PARAMETER before including to change it.  The reference swallows
 * errors; these wrappers do too (same signatures, void return), after
 * printing the status to stderr.  Callers that want the status use the
 * C-ABI entry points in gnlm.h directly.
 */
#ifndef GNLM_MAIN_H
#define GNLM_MAIN_H

#include <cstdio>
#include "gnlm.h"

#if !defined(SS) && defined(__has_include)
#if __has_include("config.h")
#include "config.h"
#endif
#endif

#ifndef GNLM_PARAMETER
#define GNLM_PARAMETER 10
#endif

class NLMeansProcessor {
public:
#if defined(SS) && defined(NS) && defined(IMAGE_SIZE_X) && defined(IMAGE_SIZE_Y) && defined(depth)
    static void NL_Means(unsigned char *GPU_input, float *GPU_result)
    {
        const int st = gnlm_nl_means_3d_u8(GPU_input, GPU_result, IMAGE_SIZE_X, IMAGE_SIZE_Y, depth, SS, NS,
                                           (float)GNLM_PARAMETER);
        if (st != GNLM_OK) std::fprintf(stderr, "NL_Means (3D): %s\n", gnlm_status_string(st));
    }
#endif
    static void NL_Means(unsigned char *GPU_input, float *GPU_result, int W, int H, int sr_size, int nb_size,
                         int image_size_x, int image_size_y)
    {
        const int st = gnlm_nl_means_2d_u8(GPU_input, GPU_result, W, H, sr_size, nb_size, image_size_x,
                                           image_size_y, (float)GNLM_PARAMETER);
        if (st != GNLM_OK) std::fprintf(stderr, "NL_Means (2D): %s\n", gnlm_status_string(st));
    }
};

#endif /* GNLM_MAIN_H */
