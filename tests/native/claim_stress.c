/*
 * CPU stress test of the direct join table's claim rule (the device code's
 * dir_word / atomicAnd claim in vh_fused.hpp build_chunk and
 * vh_pipeline.hip part_build_slice), with C11 atomics over shared 32-bit
 * words and several threads:
 *
 *   entries start all-ones; a row claims entry `off` with
 *     seen = atomic_fetch_and(word, bits | ~mask)
 *   where bits is its payload shifted into the entry and mask covers the
 *   entry; (seen & mask) == mask means it was the first row of its key, else
 *   the row is a repeat and ORs back what it saw: atomic_fetch_or(word, seen & mask).
 *
 * Checked after every round: exactly one row per key saw all-ones, and the
 * entry holds that row's payload; keys no row touched stay all-ones.
 *
 * usage: claim_stress <entry_bytes 1|2|4> <rounds> <restore 0|1>
 * restore = 0 leaves out the OR-back (a planted mistake the check must
 * catch).  Exit status 0 = every round correct, 1 = a wrong entry.
 */
#include <pthread.h>
#include <stdatomic.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define THREADS 8
#define KEYS 64          /* a few words: heavy sharing of words and keys */
#define ROWS 4096

static int g_w, g_restore;
static _Atomic uint32_t g_words[KEYS];   /* enough for KEYS entries of 4 bytes */
static uint32_t g_off[ROWS], g_pay[ROWS];
static int g_first[ROWS];
static pthread_barrier_t g_bar;

static uint64_t splitmix(uint64_t* s)
{
	uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
	return z ^ (z >> 31);
}

static void claim(int r)
{
	const uint32_t byte = g_off[r] * (uint32_t)g_w;
	const uint32_t shift = (byte & 3) * 8;
	const uint32_t mask = (g_w == 4 ? 0xFFFFFFFFu : ((1u << (8 * g_w)) - 1u)) << shift;
	_Atomic uint32_t* word = &g_words[byte >> 2];
	const uint32_t bits = (g_pay[r] << shift) & mask;
	const uint32_t seen = atomic_fetch_and(word, bits | ~mask);
	if ((seen & mask) == mask) {
		g_first[r] = 1;
	} else {
		g_first[r] = 0;
		if (g_restore)
			atomic_fetch_or(word, seen & mask);
	}
}

static void* worker(void* arg)
{
	const int t = (int)(intptr_t)arg;
	pthread_barrier_wait(&g_bar);
	for (int r = t; r < ROWS; r += THREADS)
		claim(r);
	return NULL;
}

int main(int argc, char** argv)
{
	if (argc != 4)
		return 2;
	g_w = atoi(argv[1]);
	const int rounds = atoi(argv[2]);
	g_restore = atoi(argv[3]);
	if (g_w != 1 && g_w != 2 && g_w != 4)
		return 2;
	const uint32_t empty = g_w == 4 ? 0xFFFFFFFFu : ((1u << (8 * g_w)) - 1u);
	uint64_t seed = 42;
	int bad_rounds = 0;
	for (int round = 0; round < rounds; round++) {
		const int keys = 1 + (int)(splitmix(&seed) % KEYS);
		for (int k = 0; k < KEYS; k++)
			atomic_store(&g_words[k], 0xFFFFFFFFu);
		for (int r = 0; r < ROWS; r++) {
			g_off[r] = (uint32_t)(splitmix(&seed) % (uint64_t)keys);
			/* payloads are never all-ones; many share bits with each other */
			g_pay[r] = (uint32_t)(splitmix(&seed) % (uint64_t)empty);
		}
		pthread_barrier_init(&g_bar, NULL, THREADS);
		pthread_t th[THREADS];
		for (int t = 0; t < THREADS; t++)
			pthread_create(&th[t], NULL, worker, (void*)(intptr_t)t);
		for (int t = 0; t < THREADS; t++)
			pthread_join(th[t], NULL);
		pthread_barrier_destroy(&g_bar);
		int firsts[KEYS];
		uint32_t want[KEYS];
		memset(firsts, 0, sizeof(firsts));
		for (int k = 0; k < KEYS; k++)
			want[k] = empty;
		for (int r = 0; r < ROWS; r++)
			if (g_first[r]) {
				firsts[g_off[r]]++;
				want[g_off[r]] = g_pay[r];
			}
		int bad = 0;
		for (int k = 0; k < KEYS && !bad; k++) {
			const uint32_t byte = (uint32_t)k * (uint32_t)g_w;
			const uint32_t v = (atomic_load(&g_words[byte >> 2]) >> ((byte & 3) * 8)) & empty;
			int used = 0;
			for (int r = 0; r < ROWS && !used; r++)
				used = g_off[r] == (uint32_t)k;
			if (firsts[k] != (used ? 1 : 0) || v != want[k])
				bad = 1;
		}
		bad_rounds += bad;
	}
	printf("entry_bytes=%d rounds=%d restore=%d bad_rounds=%d\n", g_w, rounds, g_restore, bad_rounds);
	return bad_rounds ? 1 : 0;
}
