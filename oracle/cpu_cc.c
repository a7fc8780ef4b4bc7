/* placeholder until the CPU path lands */
#include "oracle.h"
int orc_cpu_run(int scheme, uint64_t rows, uint32_t cols, uint64_t seed, uint64_t* table, const uint64_t* ops, const uint64_t* tx_off, uint64_t ntx, uint64_t batch_size, int threads, uint64_t* order, uint64_t* reads, orc_cpu_stats* stats) { (void)scheme;(void)rows;(void)cols;(void)seed;(void)table;(void)ops;(void)tx_off;(void)ntx;(void)batch_size;(void)threads;(void)order;(void)reads;(void)stats; return -1; }
