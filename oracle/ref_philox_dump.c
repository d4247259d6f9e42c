/* Test-only driver around the reference's Random123 Philox4x32-10
 * (RayFactorCL/Random123/philox.h), built by `make -C oracle ref` into
 * oracle/_ref/.  Reads lines "c0 c1 c2 c3 k0 k1" (hex) on stdin and prints
 * "o0 o1 o2 o3" (hex), one block per line. */
#include <stdio.h>
#include <Random123/philox.h>

int main(void)
{
    unsigned c0, c1, c2, c3, k0, k1;
    while (scanf("%x %x %x %x %x %x", &c0, &c1, &c2, &c3, &k0, &k1) == 6) {
        philox4x32_ctr_t c = {{c0, c1, c2, c3}};
        philox4x32_key_t k = {{k0, k1}};
        philox4x32_ctr_t r = philox4x32(c, k);
        printf("%08x %08x %08x %08x\n", r.v[0], r.v[1], r.v[2], r.v[3]);
    }
    return 0;
}
