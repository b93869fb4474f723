// Synthetic mesh generators and the binary mesh file used by the solver's workloads.
#pragma once
#include "gdtp_common.h"

namespace gdtp {

// Unit cube, `subdiv` x `subdiv` quads per face split in two (subdiv 16 -> 3072 triangles),
// six square charts packed 3 x 2 in a W x H atlas.
int MakeCube(int subdiv, int W, int H, TexturedMesh& m);

// Icosahedron with every face subdivided into 4^levels triangles (levels 7 -> 327,680 triangles),
// vertex radius 1 + amplitude * smooth_noise(seed), one chart per icosahedron face.
int MakeIcosphere(int levels, int W, int H, double amplitude, uint64_t seed, TexturedMesh& m);

// Torus with nu x nv quads split in two, tube radius perturbed by smooth noise, cut into
// chartsU x chartsV rectangular charts.
int MakeTorus(int nu, int nv, int chartsU, int chartsV, double R, double r, double amplitude, uint64_t seed, int W,
              int H, TexturedMesh& m);

// W x H x channels uniform [0,1) float32 values, row-major (j, i, c), from SplitMix64(seed).
void MakeNoiseTexture(int W, int H, int channels, uint64_t seed, std::vector<float>& tex);

int WriteMesh(const char* path, const TexturedMesh& m);
int ReadMesh(const char* path, TexturedMesh& m);

// PLY with per-face texcoord lists (SimpleMesh.h:118-160), ASCII or binary little-endian.
int ReadPly(const std::string& path, TexturedMesh& m, std::string& err);

}  // namespace gdtp
