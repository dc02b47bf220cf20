/*
 * Frame-serving scheduler over the MI355X renderer (sw_gpu.h).
 *
 * Requests are single views of one of several resident maps. The server keeps a FIFO of submitted
 * requests and a set of slots (each slot a context: its own HIP stream and work buffers). Whenever
 * a slot is free it forms a batch from the queue head: the head request plus later queued requests
 * of the same map and colour size, up to a view limit and as long as the batch's HBM working set
 * fits the budget left by the batches in flight (continuous batching: requests that arrive while
 * others run join the next batch as slots and memory free up). Batches of different slots run
 * concurrently on the device. Results are copied into the caller's buffers; 64-bit requests can
 * also be tonemapped on the device.
 *
 * No counterpart in the reference (its renderer draws one frame per call for one window); it is
 * the serving form of the batched path (sw_render_batch_async).
 *
 * Not thread-safe: one host thread drives a server.
 */
#ifndef SW_SERVER_H
#define SW_SERVER_H

#include "sw_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sw_server sw_server;

typedef struct sw_frame_request
{
	uint32_t map; /* index returned by sw_server_add_map */
	uint32_t color_bits; /* 32 or 64 */
	sw_view view; /* size, camera, clear flag and colour; the batch offsets are filled by the server */
	const sw_surface_desc* surfaces; /* host memory; pixels_offset relative to this view's arena */
	uint32_t num_surfaces;
	uint32_t num_lights;
	const sw_light* lights;
	const uint32_t* light_indices;
	uint32_t num_light_indices;
	uint32_t num_draws;
	const sw_draw* draws;
	const int32_t* vertex_data;
	uint64_t num_vertex_ints;
	uint64_t surface_arena_texels;
	void* framebuffer; /* optional destination: width * height ColorT elements, pitch = width */
	uint32_t* tonemapped; /* optional destination (64-bit requests): width * height Color32 */
	float exposure; /* tonemap exposure (postprocessor.rs:1005-1017), modulation 1 */
	uint32_t flags; /* SW_REQ_* */
} sw_frame_request;

/* The result buffers are pinned host memory (sw_server_host_alloc / sw_host_alloc) or device
 * memory: the DMA engine copies the results into them directly, behind the batch. Without the flag
 * they are ordinary host memory and the results pass through the slot's pinned staging memory. */
#define SW_REQ_PINNED_OUTPUT 1u

typedef struct sw_completion
{
	uint64_t ticket;
	int32_t status; /* SW_OK or an sw_status error */
	uint32_t batch_views; /* views of the batch the request ran in */
	double queue_ms; /* submit -> launch */
	double total_ms; /* submit -> result in the caller's buffer */
} sw_completion;

typedef struct sw_server_stats
{
	uint64_t submitted, completed, rejected, batches, views;
	uint64_t peak_bytes_in_flight;
	uint32_t peak_slots_busy;
	uint32_t reserved;
} sw_server_stats;

/* slots: concurrent batches (contexts / streams); hbm_budget_bytes: upper bound of the working
 * sets of the batches in flight; max_views_per_batch: batch size limit. */
int sw_server_create(int device, uint32_t slots, uint64_t hbm_budget_bytes, uint32_t max_views_per_batch, sw_server** out);
void sw_server_destroy(sw_server* server);
const char* sw_server_last_error(const sw_server* server);

/* Uploads a map (resident until the server is destroyed). */
int sw_server_add_map(sw_server* server, const sw_map_desc* desc, uint32_t* out_map);

/* Queues a request; its command buffers are copied, the result buffers must stay valid until
 * the request's completion is returned. */
int sw_server_submit(sw_server* server, const sw_frame_request* request, uint64_t* out_ticket);

/* Schedules queued requests onto free slots and collects finished batches without blocking;
 * writes up to max completions. */
int sw_server_poll(sw_server* server, sw_completion* out, uint32_t max, uint32_t* out_count);

/* Blocks until every submitted request has completed (or `max` completions are returned). */
int sw_server_drain(sw_server* server, sw_completion* out, uint32_t max, uint32_t* out_count);

/* Requests submitted and not yet returned as completions. */
uint32_t sw_server_pending(const sw_server* server);

int sw_server_get_stats(const sw_server* server, sw_server_stats* out);

/* Pinned host memory for SW_REQ_PINNED_OUTPUT result buffers (freed with sw_server_host_free, before
 * the server is destroyed). */
int sw_server_host_alloc(sw_server* server, uint64_t bytes, void** out);
void sw_server_host_free(sw_server* server, void* p);

#ifdef __cplusplus
}
#endif

#endif
