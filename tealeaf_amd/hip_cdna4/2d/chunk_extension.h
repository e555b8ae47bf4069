#pragma once

// hip_cdna4 suite (MI355X, gfx950): every field is a device pointer that the
// host never dereferences (2d/chunk.h:30-58 holds them as FieldBufferType).
typedef double* FieldBufferType;

// The fused Chebyshev and PPCG sweeps write u and sd into a second buffer
// and the glue swaps the chunk's pointer with it after each call.
typedef struct ChunkExtension
{
    double* u_alt;
    double* sd_alt;
} ChunkExtension;
