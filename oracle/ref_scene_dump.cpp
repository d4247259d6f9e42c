// Test-only driver around the reference scene loader (Scene::readScene,
// RayFactorCL/Scene.cpp:24-452), built by `make -C oracle ref` into
// oracle/_ref/.  It writes the Primitive / BoundingVolume arrays exactly as
// the reference produces them, so the product's own loader can be checked
// byte for byte.
//
// The reference leaves several fields uninitialised (Scene::useBoundingVolumes,
// util1/util2 of discs and rectangles, BV settings/child range).  Every heap
// allocation is zero-filled here, so those fields read as zero, the value the
// product loader assigns.
//
// Output (little endian): "RFSC" u32 version=1, i32 n_prims, i32 n_bvs,
// f32 globalRayDensity, i32 useBoundingVolumes, then n_prims x 112-byte
// Primitive records and n_bvs x 72-byte BoundingVolume records.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include "Scene.h"

void *operator new(std::size_t n) { void *p = std::calloc(1, n ? n : 1); if (!p) throw std::bad_alloc(); return p; }
void *operator new[](std::size_t n) { void *p = std::calloc(1, n ? n : 1); if (!p) throw std::bad_alloc(); return p; }
void operator delete(void *p) noexcept { std::free(p); }
void operator delete[](void *p) noexcept { std::free(p); }
void operator delete(void *p, std::size_t) noexcept { std::free(p); }
void operator delete[](void *p, std::size_t) noexcept { std::free(p); }

int main(int argc, char **argv)
{
    if (argc != 3) { std::fprintf(stderr, "usage: %s scene.xml out.rfscene\n", argv[0]); return 2; }
    static_assert(sizeof(Primitive) == 112, "Primitive must be 112 bytes");
    static_assert(sizeof(BoundingVolume) == 72, "BoundingVolume must be 72 bytes");
    Scene *scene = new Scene;
    scene->readScene(argv[1]);
    FILE *f = std::fopen(argv[2], "wb");
    if (!f) return 1;
    const unsigned version = 1;
    const int np = (int)scene->primitives.size(), nb = (int)scene->boundingVolumes.size();
    const int use_bv = scene->useBoundingVolumes ? 1 : 0;
    std::fwrite("RFSC", 1, 4, f);
    std::fwrite(&version, 4, 1, f);
    std::fwrite(&np, 4, 1, f);
    std::fwrite(&nb, 4, 1, f);
    std::fwrite(&scene->globalRayDensity, 4, 1, f);
    std::fwrite(&use_bv, 4, 1, f);
    if (np) std::fwrite(&scene->primitives[0], sizeof(Primitive), np, f);
    if (nb) std::fwrite(&scene->boundingVolumes[0], sizeof(BoundingVolume), nb, f);
    std::fclose(f);
    return 0;
}
