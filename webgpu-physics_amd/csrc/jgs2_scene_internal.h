// Host scene representation shared by the scene builders of libjgs2hip.
#pragma once

#include <cstdint>
#include <vector>

#include "jgs2_internal.h"
#include "jgs2hip_scene.h"

struct jgs2_scene {
  uint32_t vertex_count = 0;
  uint32_t tet_count = 0;
  uint32_t cubature_k = 0;
  std::vector<float> positions, vertex_rest;
  std::vector<uint32_t> vertex_info, tet_indices, adjacency;
  std::vector<float> tet_inverse_dm, tet_meta, tet_rest_stiffness;
  std::vector<uint32_t> cubature_tet_ids;
  std::vector<float> cubature_weights, cubature_basis;
  std::vector<float> forces, targets, stiffnesses;
  // Optional cloth layer.
  bool has_cloth = false;
  std::vector<uint32_t> cloth_triangles, cloth_hinges;
  std::vector<float> cloth_inverse, cloth_areas, hinge_angles, hinge_lengths;
  jgs2_cloth_input cloth{};
  // Boundary surface and optional static IPC candidates.
  std::vector<uint32_t> surface_triangles, surface_edges;
  bool has_contact = false;
  std::vector<uint32_t> contact_indices;
  jgs2_contact_input contact{};
};

namespace jgs2 {

// A tetrahedral mesh before packing (TetrahedralMesh, src/simulation/cpu/
// types.ts): f64 positions, tetrahedra, one material per tetrahedron, fixed
// flags and body ids per vertex.
struct Mesh {
  std::vector<double> x;
  std::vector<uint32_t> tets;
  std::vector<uint32_t> material;
  std::vector<uint8_t> fixed;
  std::vector<uint32_t> body;
  uint32_t vertex_count() const { return uint32_t(fixed.size()); }
  uint32_t tet_count() const { return uint32_t(material.size()); }
};

struct Material {
  int32_t model;  // JGS2_MATERIAL_*
  double density, young, poisson;
};

// generateRegularCuboidMesh without orientation (mesh.ts:101-191): grid
// vertices in x-fastest order and six Kuhn tetrahedra per cell.
void cuboid_grid(const uint32_t cells[3], const double origin[3], const double size[3], uint32_t material,
                 uint32_t body, bool fixed, Mesh& m);
// orientTetrahedra (mesh.ts:80-99): swap b and c of negatively oriented
// tetrahedra; fails on degenerate ones.
jgs2_status orient(Mesh& m);
// appendTetrahedralMeshes (mesh.ts:284-326).
void append(Mesh& into, const Mesh& m);
// Rest data, materials, lumped masses, canonical CSR, vertex records and the
// cubature record generator (toJGS2GpuInput, src/scenes/index.ts:975-1109).
// `cubature_h` is the length scale of the generated cubature bases.
jgs2_status pack_scene(const Mesh& m, const std::vector<Material>& materials, uint32_t cubature_k,
                       bool rest_stiffness, double cubature_h, jgs2_scene* s);
// extractBoundarySurface (mesh.ts:366-440) of the scene's tetrahedra.
jgs2_status boundary_surface(const Mesh& m, std::vector<uint32_t>& triangles, std::vector<uint32_t>& edges);

}  // namespace jgs2
