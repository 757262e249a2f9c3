/* mtrand.c -- the reference's Brownian random stream on the host: the
 * Mersenne Twister MT19937 (Matsumoto & Nishimura, ACM TOMACS 8, 3-30,
 * 1998: 624 words of state, tempering, genrand_real2 = int32 / 2^32 on
 * [0, 1)) and Marsaglia's polar method returning x r first and saving
 * y r, the order of KIrand_Gaussian (KIrand.c:108-135).  Seeded as
 * KIrand_init_genrand (KIrand.c:137-157). */
#include <math.h>
#include "stokes_internal.h"

#define MT_N 624
#define MT_M 397

void
stokes_hip_mt_seed (stokes_hip_mt *r, unsigned long seed)
{
  r->mt[0] = seed & 0xffffffffUL;
  for (int i = 1; i < MT_N; i++)
    r->mt[i] = (1812433253UL * (r->mt[i - 1] ^ (r->mt[i - 1] >> 30)) + (unsigned long) i)
               & 0xffffffffUL;
  r->idx = MT_N;
  r->has_saved = 0;
  r->saved = 0.0;
}

static unsigned long
mt_next (stokes_hip_mt *r)
{
  if (r->idx >= MT_N)
    {
      /* regenerate the whole state: y = upper bit of word k, lower 31
       * bits of word k + 1; word k = word k + M  xor  (y >> 1)  xor
       * (0x9908b0df if y is odd) */
      for (int k = 0; k < MT_N; k++)
        {
          const unsigned long y = (r->mt[k] & 0x80000000UL)
                                  | (r->mt[(k + 1) % MT_N] & 0x7fffffffUL);
          r->mt[k] = r->mt[(k + MT_M) % MT_N] ^ (y >> 1) ^ ((y & 1UL) ? 0x9908b0dfUL : 0UL);
        }
      r->idx = 0;
    }
  unsigned long y = r->mt[r->idx++];
  y ^= y >> 11;
  y ^= (y << 7) & 0x9d2c5680UL;
  y ^= (y << 15) & 0xefc60000UL;
  y ^= y >> 18;
  return y & 0xffffffffUL;
}

double
stokes_hip_mt_gaussian (stokes_hip_mt *r)
{
  if (r->has_saved)
    {
      r->has_saved = 0;
      return r->saved;
    }
  double x, y, q;
  do
    {
      x = 2.0 * ((double) mt_next (r) * (1.0 / 4294967296.0)) - 1.0;
      y = 2.0 * ((double) mt_next (r) * (1.0 / 4294967296.0)) - 1.0;
      q = x * x + y * y;
    }
  while (q > 1.0 || q == 0.0);
  q = sqrt (-2.0 * log (q) / q);
  r->saved = y * q;
  r->has_saved = 1;
  return x * q;
}
