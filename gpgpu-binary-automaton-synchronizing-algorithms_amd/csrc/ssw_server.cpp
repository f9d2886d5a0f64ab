// ssw_server.cpp -- placeholder until the serving loop lands.
#include "ssw_host.h"
extern "C" {
int ssw_server_create(ssw_context *, const ssw_server_config *, ssw_server **) { return SSW_ERR_INTERNAL; }
int ssw_server_submit(ssw_server *, int32_t, const int32_t *, const int32_t *, uint64_t *) { return SSW_ERR_INTERNAL; }
int ssw_server_poll(ssw_server *, uint64_t, ssw_result *) { return SSW_ERR_INTERNAL; }
int ssw_server_wait(ssw_server *, uint64_t, int32_t, ssw_result *) { return SSW_ERR_INTERNAL; }
int ssw_server_stats_get(ssw_server *, ssw_server_stats *) { return SSW_ERR_INTERNAL; }
int ssw_server_destroy(ssw_server *) { return SSW_ERR_INTERNAL; }
}
