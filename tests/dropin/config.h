/* Geometry for the drop-in check, in the reference's config.h form
 * (3DNLM/config.h:4-9): SS = search window size, NS = patch size. */
#ifndef CONFIG_H
#define CONFIG_H
#define SS 7
#define NS 3
#define IMAGE_SIZE_X 40
#define IMAGE_SIZE_Y 36
#define depth 12
#endif
