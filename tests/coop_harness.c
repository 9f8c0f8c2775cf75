/* CPU build of the wave-cooperative network programs (generated by
 * warp-reactor_amd/gen/coop.py) next to the plain network code (gen/emit.py),
 * both with this build's libm (wr_math.h): every vector statement of the
 * cooperative code runs as a loop over the 64 lanes of a wave, on a byte
 * array standing in for the wave's LDS region.  Test infrastructure:
 * tests/test_coop.py compiles and calls it. */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "wr_math.h"

#define CHEM_EXP(x) wr_exp(x)
#define CHEM_LOG(x) wr_log(x)

/* plain network: chem_fused / chem_rhs */
#define CHEM_FN static
#define CHEM_SINK_PARAMS , double *jac
#define CHEM_JAC_STORE(r, c, v) (jac[(r) * 15 + (c)] = (v))
#include "chem_network.h"

static char coop_smem[1 << 18];

static inline double coop_flip(double x, uint32_t w) {
    return wr_from_bits(wr_bits_of(x) ^ ((uint64_t)(w & 0x80000000u) << 32));
}
static inline double coop_sgn(double x, uint32_t w) {
    uint64_t b = wr_bits_of(x);
    if (w & 0x40000000u) b &= ~(1ull << 63);
    return wr_from_bits(b ^ ((uint64_t)(w & 0x80000000u) << 32));
}

#define COOP_FN static
#define COOP_PARAMS void
#define COOP_DATA static const
#define COOP_DU(x) static uint32_t x[64]
#define COOP_DD(x) static double x[64]
#define COOP_V(x) x[ln]
#define COOP_L(...)                       \
    for (int ln = 0; ln < 64; ++ln) {     \
        __VA_ARGS__                       \
    }
#define COOP_TW(tab, row) (tab)[(row) * 64 + ln]
#ifdef COOP_TYPE_CHECK
/* every access records its type (1 double, 2 boolean word) on its 8-byte slot;
 * an access of the other type than the slot's last one is counted */
static uint32_t coop_typed(uint32_t a, unsigned char t);
#define COOP_LD(a) (*(const double *)(coop_smem + coop_typed((a), 1)))
#define COOP_LDU(a) (*(const uint32_t *)(coop_smem + coop_typed((a), 2)))
#define COOP_ST(a, v) (*(double *)(coop_smem + coop_typed((a), 1)) = (v))
#define COOP_STU(a, v) (*(uint32_t *)(coop_smem + coop_typed((a), 2)) = (v))
#else
#define COOP_LD(a) (*(const double *)(coop_smem + (a)))
#define COOP_LDU(a) (*(const uint32_t *)(coop_smem + (a)))
#define COOP_ST(a, v) (*(double *)(coop_smem + (a)) = (v))
#define COOP_STU(a, v) (*(uint32_t *)(coop_smem + (a)) = (v))
#endif
#include "coop_network.h"

#ifdef COOP_TYPE_CHECK
static unsigned char coop_tag[(1 << 18) / 8];
static long coop_conflicts;
static uint32_t coop_typed(uint32_t a, unsigned char t) {
    /* the trash slot takes the results of idle lanes of every class; the
     * constants (which idle lanes of every class read) are written before a
     * program runs and never stored by one */
    if (a > COOP_OFF_TRASH) {
        unsigned char *g = &coop_tag[a / 8];
        if (*g && *g != t) ++coop_conflicts;
        *g = t;
    }
    return a;
}
/* accesses of one 8-byte LDS slot with two types since the library loaded */
long coop_type_conflicts(void) { return coop_conflicts; }
#endif

static void coop_load_inputs(double T, const double *X) {
    memcpy(coop_smem + COOP_OFF_CONST, coop_const_bits, sizeof(coop_const_bits));
    memcpy(coop_smem + COOP_OFF_IN, &T, 8);
    memcpy(coop_smem + COOP_OFF_IN + 8, X, 14 * 8);
}

int coop_lds_bytes(void) { return COOP_LDS_BYTES; }

void coop_eval_fused(double T, const double *X, double *f, double *J) {
    coop_load_inputs(T, X);
    coop_fused();
    memcpy(f, coop_smem + COOP_OFF_VEC + 15 * 8, 15 * 8);
    memcpy(J, coop_smem + COOP_OFF_J, 225 * 8);
}

void coop_eval_rhs(double T, const double *X, double *f) {
    coop_load_inputs(T, X);
    coop_rhs();
    memcpy(f, coop_smem + COOP_OFF_VEC + 90 * 8, 15 * 8);
}

void plain_eval_fused(double T, const double *X, double *f, double *J) { chem_fused(T, X, f, J); }
void plain_eval_rhs(double T, const double *X, double *f) { chem_rhs(T, X, f); }

/* debugging aid: the LDS image after the last evaluation */
const char *coop_lds_image(void) { return coop_smem; }
