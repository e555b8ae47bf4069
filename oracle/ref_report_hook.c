/*
 * Report hook for the reference host binaries built under oracle/_ref/
 * (test infrastructure).  Compiled with -DDIFFUSE_OVERLOAD, which makes the
 * reference main call diffuse_overload() instead of diffuse()
 * (2d/main.c:31-36, 3d/main.c:30-35).  It runs the reference's own diffuse()
 * unchanged and then prints all four field_summary values, which the
 * reference computes but only partly prints (2d/drivers/field_summary_driver.c:40-41).
 * The same hook is linked with the omp3 suite and with the hip_cdna4 suite.
 */
#include <stdio.h>
#include "application.h"
#include "kernel_interface.h"

void diffuse_overload(Chunk* chunks, Settings* settings)
{
    diffuse(chunks, settings);
    double vol = 0.0, mass = 0.0, ie = 0.0, temp = 0.0;
    run_field_summary(&chunks[0], settings, &vol, &mass, &ie, &temp);
    printf("TL_FIELD_SUMMARY %.17e %.17e %.17e %.17e\n", vol, mass, ie, temp);
    fflush(stdout);
}
