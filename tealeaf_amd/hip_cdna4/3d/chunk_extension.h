#pragma once

// hip_cdna4 suite (MI355X, gfx950): the 3D chunk keeps plain double* fields
// (3d/chunk.h:45-91); they are device pointers the host never dereferences.
// u_alt / sd_alt are the ping-pong partners of vec_u / vec_sd used by the
// fused Chebyshev and PPCG sweeps.
typedef struct chunk_extension_t
{
    double* u_alt;
    double* sd_alt;
} chunk_extension_t;
