// Frame-serving scheduler (include/sw_server.h): continuous batching of single-view requests onto
// several contexts of one device, admission by an HBM budget. Built on the public C ABI only.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/sw_server.h"

namespace
{

typedef std::chrono::steady_clock Clock;

double ms_between(Clock::time_point a, Clock::time_point b) { return std::chrono::duration<double, std::milli>(b - a).count(); }

struct Request
{
	uint64_t ticket = 0;
	sw_frame_request r{};
	std::vector<sw_surface_desc> surfaces;
	std::vector<sw_light> lights;
	std::vector<uint32_t> light_indices;
	std::vector<sw_draw> draws;
	std::vector<int32_t> vdata;
	uint64_t bytes = 0; // HBM working set estimate
	Clock::time_point t_submit;
	Clock::time_point t_launch;
	uint64_t fb_offset = 0; // in the batch output, ColorT elements
	uint64_t stage_fb = 0, stage_ldr = 0; // byte offsets of its results in the slot's pinned staging
	uint64_t ldr_offset = 0; // in the slot's tonemap output, Color32 elements
};

struct Slot
{
	sw_ctx* ctx = nullptr;
	bool busy = false;
	std::vector<Request> reqs;
	void* out_dev = nullptr;
	uint64_t out_cap = 0;
	void* ldr_dev = nullptr;
	uint64_t ldr_cap = 0;
	void* staging = nullptr; // pinned host memory the results are copied into behind the batch
	uint64_t staging_cap = 0;
	uint64_t bytes = 0;
};

uint64_t request_bytes(const sw_frame_request& r)
{
	const uint64_t c = r.color_bits == 64 ? 8 : 4;
	const uint64_t px = uint64_t(r.view.width) * r.view.height;
	uint64_t b = px * c + r.surface_arena_texels * c;
	b += uint64_t(r.num_draws) * (uint64_t(r.view.height) * 32 + 24 * 48 + 48 + sizeof(sw_draw));
	b += uint64_t(r.num_surfaces) * (sizeof(sw_surface_desc) + 16) + r.num_vertex_ints * 4;
	b += uint64_t(r.num_lights) * sizeof(sw_light) + uint64_t(r.num_light_indices) * 4;
	if(r.tonemapped && r.color_bits == 64)
		b += px * 4;
	return b;
}

} // namespace

struct sw_server
{
	int device = 0;
	uint64_t budget = 0;
	uint32_t max_views = 1;
	std::string error;
	std::deque<Request> queue;
	std::vector<Slot> slots;
	std::vector<sw_map*> maps;
	uint64_t in_flight_bytes = 0;
	uint64_t next_ticket = 1;
	uint32_t outstanding = 0; // submitted, completion not yet returned
	std::vector<sw_completion> ready;
	sw_server_stats stats{};
};

namespace
{

int fail(sw_server* s, int code, const std::string& msg)
{
	s->error = msg;
	return code;
}

void complete(sw_server* s, const Request& q, int status, uint32_t batch_views, Clock::time_point now)
{
	sw_completion c;
	c.ticket = q.ticket;
	c.status = status;
	c.batch_views = batch_views;
	c.queue_ms = status == SW_OK || batch_views ? ms_between(q.t_submit, q.t_launch) : ms_between(q.t_submit, now);
	c.total_ms = ms_between(q.t_submit, now);
	s->ready.push_back(c);
	if(status == SW_OK)
		s->stats.completed++;
	else
		s->stats.rejected++;
}

int ensure_dev(sw_ctx* ctx, void*& p, uint64_t& cap, uint64_t bytes)
{
	if(bytes <= cap && p)
		return SW_OK;
	if(p)
		sw_device_free(ctx, p);
	p = nullptr;
	cap = 0;
	const uint64_t want = std::max<uint64_t>(bytes + bytes / 4, 1 << 20);
	int r = sw_device_alloc(ctx, want, &p);
	if(r == SW_OK)
		cap = want;
	return r;
}

// Queues behind the batch: the tonemap of each 64-bit request that asks for one and the copies of
// every result into the slot's pinned staging memory.
int queue_results(Slot& slot)
{
	const float mod[3] = {1.0f, 1.0f, 1.0f};
	int st = SW_OK;
	for(Request& q : slot.reqs)
	{
		const uint64_t c = q.r.color_bits == 64 ? 8 : 4;
		const uint32_t w = q.r.view.width, h = q.r.view.height;
		const bool direct = (q.r.flags & SW_REQ_PINNED_OUTPUT) != 0;
		const char* src = static_cast<const char*>(slot.out_dev) + q.fb_offset * c;
		if(st == SW_OK && q.r.framebuffer)
			st = sw_copy_from_device_async(slot.ctx,
				direct ? q.r.framebuffer : static_cast<char*>(slot.staging) + q.stage_fb, src, uint64_t(w) * h * c);
		if(st == SW_OK && q.r.tonemapped && q.r.color_bits == 64)
		{
			uint32_t* ldr = static_cast<uint32_t*>(slot.ldr_dev) + q.ldr_offset;
			st = sw_tonemap_device(slot.ctx, src, w, h, q.r.exposure, mod, ldr, size_t(w) * 4, nullptr);
			if(st == SW_OK)
				st = sw_copy_from_device_async(slot.ctx,
					direct ? static_cast<void*>(q.r.tonemapped) : static_cast<char*>(slot.staging) + q.stage_ldr, ldr,
					uint64_t(w) * h * 4);
		}
	}
	return st;
}

// A finished batch: status, results into the callers' buffers, budget back.
void collect(sw_server* s, Slot& slot)
{
	int status = sw_batch_wait(slot.ctx);
	if(status == SW_OK)
	{
		if(sw_last_reruns(slot.ctx))
		{
			// the raster stage ran again after the copies were queued: copy again
			status = queue_results(slot);
			if(status == SW_OK)
				status = sw_sync(slot.ctx);
		}
	}
	const uint32_t views = uint32_t(slot.reqs.size());
	const Clock::time_point now = Clock::now();
	for(Request& q : slot.reqs)
	{
		const uint64_t c = q.r.color_bits == 64 ? 8 : 4;
		const uint64_t px = uint64_t(q.r.view.width) * q.r.view.height;
		const bool staged = status == SW_OK && !(q.r.flags & SW_REQ_PINNED_OUTPUT);
		if(staged && q.r.framebuffer)
			std::memcpy(q.r.framebuffer, static_cast<const char*>(slot.staging) + q.stage_fb, px * c);
		if(staged && q.r.tonemapped && q.r.color_bits == 64)
			std::memcpy(q.r.tonemapped, static_cast<const char*>(slot.staging) + q.stage_ldr, px * 4);
		complete(s, q, status, views, now);
	}
	slot.reqs.clear();
	s->in_flight_bytes -= slot.bytes;
	slot.bytes = 0;
	slot.busy = false;
}

int ensure_host(sw_ctx* ctx, void*& p, uint64_t& cap, uint64_t bytes)
{
	if(bytes <= cap && p)
		return SW_OK;
	if(p)
		sw_host_free(ctx, p);
	p = nullptr;
	cap = 0;
	const uint64_t want = std::max<uint64_t>(bytes + bytes / 4, 1 << 20);
	int r = sw_host_alloc(ctx, want, &p);
	if(r == SW_OK)
		cap = want;
	return r;
}

// Forms a batch from the queue head for an idle slot and launches it. Returns false when nothing
// could be launched (empty queue or the head waits for budget).
bool launch_next(sw_server* s, Slot& slot)
{
	while(!s->queue.empty())
	{
		Request& head = s->queue.front();
		if(head.bytes > s->budget)
		{
			// can never fit: reject
			complete(s, head, SW_ERR_BUDGET, 0, Clock::now());
			s->queue.pop_front();
			continue;
		}
		break;
	}
	if(s->queue.empty())
		return false;
	const uint64_t free_bytes = s->budget - s->in_flight_bytes;
	if(s->queue.front().bytes > free_bytes)
		return false;
	const uint32_t map = s->queue.front().r.map, bits = s->queue.front().r.color_bits;
	std::vector<Request>& take = slot.reqs;
	take.clear();
	uint64_t bytes = 0;
	const size_t scan = std::min<size_t>(s->queue.size(), 1024);
	std::vector<size_t> picked;
	for(size_t i = 0; i < scan && take.size() < s->max_views; ++i)
	{
		const Request& q = s->queue[i];
		if(q.r.map != map || q.r.color_bits != bits || bytes + q.bytes > free_bytes)
			continue;
		bytes += q.bytes;
		picked.push_back(i);
		take.push_back(std::move(s->queue[i]));
	}
	for(size_t k = picked.size(); k-- > 0;)
		s->queue.erase(s->queue.begin() + std::ptrdiff_t(picked[k]));

	// concatenate the views' command buffers into one batch
	std::vector<sw_view> views(take.size());
	std::vector<sw_surface_desc> surfaces;
	std::vector<sw_light> lights;
	std::vector<uint32_t> light_indices;
	std::vector<sw_draw> draws;
	std::vector<int32_t> vdata;
	uint64_t fb = 0, arena = 0, stage = 0, ldr = 0;
	const Clock::time_point now = Clock::now();
	for(size_t i = 0; i < take.size(); ++i)
	{
		Request& q = take[i];
		q.t_launch = now;
		q.fb_offset = fb;
		const uint64_t px = uint64_t(q.r.view.width) * q.r.view.height;
		const bool staged = !(q.r.flags & SW_REQ_PINNED_OUTPUT);
		q.stage_fb = stage;
		if(staged && q.r.framebuffer)
			stage += px * (bits == 64 ? 8 : 4);
		q.stage_ldr = stage;
		q.ldr_offset = ldr;
		if(q.r.tonemapped && bits == 64)
		{
			if(staged)
				stage += px * 4;
			ldr += px;
		}
		sw_view& v = views[i];
		v = q.r.view;
		v.first_surface = uint32_t(surfaces.size());
		v.num_surfaces = uint32_t(q.surfaces.size());
		v.first_draw = uint32_t(draws.size());
		v.num_draws = uint32_t(q.draws.size());
		v.fb_offset = fb;
		v.surface_arena_offset = arena;
		const uint32_t light_base = uint32_t(lights.size()), index_base = uint32_t(light_indices.size());
		const uint32_t vbase = uint32_t(vdata.size());
		for(sw_surface_desc d : q.surfaces)
		{
			d.first_light += index_base;
			surfaces.push_back(d);
		}
		lights.insert(lights.end(), q.lights.begin(), q.lights.end());
		for(uint32_t li : q.light_indices)
			light_indices.push_back(li + light_base);
		for(sw_draw d : q.draws)
		{
			d.first_vertex_int += vbase;
			draws.push_back(d);
		}
		vdata.insert(vdata.end(), q.vdata.begin(), q.vdata.end());
		fb += uint64_t(v.width) * v.height;
		arena += q.r.surface_arena_texels;
	}
	const uint64_t c = bits == 64 ? 8 : 4;
	int r = ensure_dev(slot.ctx, slot.out_dev, slot.out_cap, std::max<uint64_t>(fb, 1) * c);
	if(r == SW_OK)
		r = ensure_dev(slot.ctx, slot.ldr_dev, slot.ldr_cap, std::max<uint64_t>(ldr, 1) * 4);
	if(r == SW_OK)
		r = ensure_host(slot.ctx, slot.staging, slot.staging_cap, std::max<uint64_t>(stage, 1));
	if(r == SW_OK)
	{
		sw_batch b{};
		b.views = views.data();
		b.num_views = uint32_t(views.size());
		b.color_bits = bits;
		b.surfaces = surfaces.data();
		b.num_surfaces = uint32_t(surfaces.size());
		b.num_lights = uint32_t(lights.size());
		b.lights = lights.data();
		b.light_indices = light_indices.data();
		b.num_light_indices = uint32_t(light_indices.size());
		b.num_draws = uint32_t(draws.size());
		b.draws = draws.data();
		b.vertex_data = vdata.data();
		b.num_vertex_ints = vdata.size();
		b.surface_arena_texels = arena;
		b.framebuffer_texels = fb;
		b.inputs_on_device = 0;
		r = sw_render_batch_async(slot.ctx, s->maps[map], &b, slot.out_dev);
		if(r == SW_OK)
			r = queue_results(slot);
		if(r != SW_OK)
			(void)sw_batch_wait(slot.ctx); // leave the context idle
	}
	if(r != SW_OK)
	{
		s->error = sw_last_error(slot.ctx);
		for(const Request& q : take)
			complete(s, q, r, uint32_t(take.size()), Clock::now());
		take.clear();
		return true;
	}
	slot.busy = true;
	slot.bytes = bytes;
	s->in_flight_bytes += bytes;
	s->stats.batches++;
	s->stats.views += take.size();
	s->stats.peak_bytes_in_flight = std::max(s->stats.peak_bytes_in_flight, s->in_flight_bytes);
	uint32_t busy = 0;
	for(const Slot& sl : s->slots)
		busy += sl.busy ? 1u : 0u;
	s->stats.peak_slots_busy = std::max(s->stats.peak_slots_busy, busy);
	return true;
}

void step(sw_server* s)
{
	for(Slot& slot : s->slots)
		if(slot.busy && sw_batch_poll(slot.ctx) != 0)
			collect(s, slot);
	for(Slot& slot : s->slots)
		if(!slot.busy && !launch_next(s, slot))
			break;
}

uint32_t take_ready(sw_server* s, sw_completion* out, uint32_t max)
{
	const uint32_t n = uint32_t(std::min<size_t>(max, s->ready.size()));
	for(uint32_t i = 0; i < n; ++i)
		out[i] = s->ready[i];
	s->ready.erase(s->ready.begin(), s->ready.begin() + n);
	s->outstanding -= n;
	return n;
}

} // namespace

extern "C" {

int sw_server_create(int device, uint32_t slots, uint64_t hbm_budget_bytes, uint32_t max_views_per_batch, sw_server** out)
{
	if(!out || slots == 0 || slots > 16 || max_views_per_batch == 0 || hbm_budget_bytes == 0)
		return SW_ERR_INVALID_ARGUMENT;
	*out = nullptr;
	sw_server* s = new(std::nothrow) sw_server();
	if(!s)
		return SW_ERR_OUT_OF_MEMORY;
	s->device = device;
	s->budget = hbm_budget_bytes;
	s->max_views = max_views_per_batch;
	s->slots.resize(slots);
	for(Slot& slot : s->slots)
	{
		const int r = sw_ctx_create(device, &slot.ctx);
		if(r != SW_OK)
		{
			sw_server_destroy(s);
			return r;
		}
	}
	*out = s;
	return SW_OK;
}

void sw_server_destroy(sw_server* s)
{
	if(!s)
		return;
	for(Slot& slot : s->slots)
		if(slot.ctx)
			(void)sw_batch_wait(slot.ctx);
	sw_ctx* owner = s->slots.empty() ? nullptr : s->slots[0].ctx;
	for(sw_map* m : s->maps)
		sw_map_release(owner, m);
	for(Slot& slot : s->slots)
		if(slot.ctx)
		{
			if(slot.out_dev)
				sw_device_free(slot.ctx, slot.out_dev);
			if(slot.ldr_dev)
				sw_device_free(slot.ctx, slot.ldr_dev);
			if(slot.staging)
				sw_host_free(slot.ctx, slot.staging);
			sw_ctx_destroy(slot.ctx);
		}
	delete s;
}

const char* sw_server_last_error(const sw_server* s) { return s ? s->error.c_str() : "null server"; }

int sw_server_host_alloc(sw_server* s, uint64_t bytes, void** out)
{
	if(!s || !out)
		return SW_ERR_INVALID_ARGUMENT;
	const int r = sw_host_alloc(s->slots[0].ctx, bytes, out);
	if(r != SW_OK)
		return fail(s, r, sw_last_error(s->slots[0].ctx));
	return SW_OK;
}

void sw_server_host_free(sw_server* s, void* p)
{
	if(s && p)
		sw_host_free(s->slots[0].ctx, p);
}

int sw_server_add_map(sw_server* s, const sw_map_desc* desc, uint32_t* out_map)
{
	if(!s || !desc || !out_map)
		return s ? fail(s, SW_ERR_INVALID_ARGUMENT, "null argument") : SW_ERR_INVALID_ARGUMENT;
	sw_map* m = nullptr;
	const int r = sw_map_upload(s->slots[0].ctx, desc, &m);
	if(r != SW_OK)
		return fail(s, r, sw_last_error(s->slots[0].ctx));
	*out_map = uint32_t(s->maps.size());
	s->maps.push_back(m);
	return SW_OK;
}

int sw_server_submit(sw_server* s, const sw_frame_request* r, uint64_t* out_ticket)
{
	if(!s || !r)
		return s ? fail(s, SW_ERR_INVALID_ARGUMENT, "null argument") : SW_ERR_INVALID_ARGUMENT;
	if(r->map >= s->maps.size())
		return fail(s, SW_ERR_INVALID_ARGUMENT, "unknown map");
	if(r->color_bits != 32 && r->color_bits != 64)
		return fail(s, SW_ERR_INVALID_ARGUMENT, "color_bits must be 32 or 64");
	if((r->num_surfaces && !r->surfaces) || (r->num_draws && !r->draws) || (r->num_vertex_ints && !r->vertex_data) ||
		(r->num_lights && !r->lights) || (r->num_light_indices && !r->light_indices))
		return fail(s, SW_ERR_INVALID_ARGUMENT, "null command buffer");
	if(r->flags & ~SW_REQ_PINNED_OUTPUT)
		return fail(s, SW_ERR_INVALID_ARGUMENT, "unknown request flags");
	Request q;
	q.ticket = s->next_ticket++;
	q.r = *r;
	q.surfaces.assign(r->surfaces, r->surfaces + r->num_surfaces);
	q.lights.assign(r->lights, r->lights + r->num_lights);
	q.light_indices.assign(r->light_indices, r->light_indices + r->num_light_indices);
	q.draws.assign(r->draws, r->draws + r->num_draws);
	q.vdata.assign(r->vertex_data, r->vertex_data + r->num_vertex_ints);
	q.bytes = request_bytes(*r);
	q.t_submit = Clock::now();
	s->queue.push_back(std::move(q));
	s->outstanding++;
	s->stats.submitted++;
	if(out_ticket)
		*out_ticket = s->next_ticket - 1;
	return SW_OK;
}

int sw_server_poll(sw_server* s, sw_completion* out, uint32_t max, uint32_t* out_count)
{
	if(!s || (!out && max))
		return SW_ERR_INVALID_ARGUMENT;
	step(s);
	const uint32_t n = take_ready(s, out, max);
	if(out_count)
		*out_count = n;
	return SW_OK;
}

int sw_server_drain(sw_server* s, sw_completion* out, uint32_t max, uint32_t* out_count)
{
	if(!s || (!out && max))
		return SW_ERR_INVALID_ARGUMENT;
	uint32_t n = 0;
	for(;;)
	{
		step(s);
		n += take_ready(s, out + n, max - n);
		if(s->outstanding == 0 || n == max)
			break;
		bool busy = false;
		for(const Slot& slot : s->slots)
			busy = busy || slot.busy;
		if(!busy && s->queue.empty() && s->ready.empty())
			break;
		std::this_thread::yield();
	}
	if(out_count)
		*out_count = n;
	return SW_OK;
}

uint32_t sw_server_pending(const sw_server* s) { return s ? s->outstanding : 0; }

int sw_server_get_stats(const sw_server* s, sw_server_stats* out)
{
	if(!s || !out)
		return SW_ERR_INVALID_ARGUMENT;
	*out = s->stats;
	return SW_OK;
}

} // extern "C"
