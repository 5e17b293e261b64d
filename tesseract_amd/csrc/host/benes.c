/* Beneš routing over the 9-cube (reference src/benes.c): 17 rounds across dims 0,1..8..1,0; in
   round s machine k takes the value across benes_dim(s) when ctrl[s][k] = 1, and both machines of
   a pair always carry the same setting.

   The looping algorithm, written as a recursion over sub-networks: the values entering a
   sub-network of 2^(9-d) positions (positions that agree in bits 0..d-1) are split into two halves
   by bit d; the two values of every input pair and of every output pair must take different
   halves. Walking the cycles those constraints form assigns every value a half; each half is then
   routed the same way one dimension up. */
#include <string.h>

#include "tesseract_host.h"

typedef struct {
    uint16_t from;  /* position on the way in */
    uint16_t to;    /* position on the way out */
} route_item;

static void route(route_item *v, int n, int d, uint8_t ctrl[TS_BENES_ROUNDS][TS_CUBE_SIZE])
{
    const unsigned bit = 1u << d;
    const int in_round = d, out_round = TS_BENES_ROUNDS - 1 - d;
    if (d == TS_CUBE_DIMS - 1) {
        for (int i = 0; i < n; i++)
            ctrl[in_round][v[i].from] = ((v[i].from ^ v[i].to) & bit) != 0;
        return;
    }
    /* index of the value at each entry / exit position (positions are < 512) */
    static __thread int at_from[TS_CUBE_DIMS][TS_CUBE_SIZE], at_to[TS_CUBE_DIMS][TS_CUBE_SIZE];
    int *by_from = at_from[d], *by_to = at_to[d];
    signed char half[TS_CUBE_SIZE];
    for (int i = 0; i < n; i++) {
        by_from[v[i].from] = i;
        by_to[v[i].to] = i;
        half[i] = -1;
    }
    for (int start = 0; start < n; start++) {
        int cur = start;
        while (half[cur] < 0) {
            half[cur] = 0;
            int partner_out = by_to[v[cur].to ^ bit];     /* shares cur's exit pair */
            half[partner_out] = 1;
            cur = by_from[v[partner_out].from ^ bit];      /* shares that one's entry pair */
        }
    }
    route_item sub[TS_CUBE_SIZE];
    int fill[2] = {0, n / 2};
    for (int i = 0; i < n; i++) {
        unsigned h = (unsigned)half[i];
        uint16_t from = (uint16_t)((v[i].from & ~bit) | (h << d));
        uint16_t to = (uint16_t)((v[i].to & ~bit) | (h << d));
        ctrl[in_round][v[i].from] = from != v[i].from;
        ctrl[out_round][to] = to != v[i].to;
        sub[fill[h]].from = from;
        sub[fill[h]++].to = to;
    }
    route(sub, n / 2, d + 1, ctrl);
    route(sub + n / 2, n / 2, d + 1, ctrl);
}

/* the dimension round s crosses: 0, 1, .., 8, .., 1, 0 */
int ts_benes_dim(int round)
{
    return round < TS_CUBE_DIMS ? round : TS_BENES_ROUNDS - 1 - round;
}

void ts_benes_route(const uint16_t *dest, uint8_t ctrl[TS_BENES_ROUNDS][TS_CUBE_SIZE])
{
    route_item v[TS_CUBE_SIZE];
    for (int k = 0; k < TS_CUBE_SIZE; k++) {
        v[k].from = (uint16_t)k;
        v[k].to = dest[k];
    }
    memset(ctrl, 0, TS_BENES_ROUNDS * TS_CUBE_SIZE);
    route(v, TS_CUBE_SIZE, 0, ctrl);
}
