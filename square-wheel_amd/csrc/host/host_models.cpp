// Dynamic lights, triangle models and decals of a frame.
#include <algorithm>
#include <cstring>
#include <random>

#include "host_frame_state.h"

using namespace swh;

namespace swh
{

void build_dynamic_lights(swh_frame& f, const swh_map& m, const swh_frame_config& cfg)
{
	f.lights.clear();
	f.leaf_lights.clear();
	if(cfg.num_dynamic_lights == 0 || m.empty_cells.empty())
		return;
	f.leaf_lights.assign(m.leafs.size(), std::vector<uint32_t>());
	std::mt19937_64 g(uint64_t(cfg.seed) * 0x2545F4914F6CDD1Dull + 17);
	auto uni = [&](float a, float b) { return a + (b - a) * float(double(g() >> 11) * (1.0 / 9007199254740992.0)); };
	const float CELL = 128.0f;
	for(uint32_t i = 0; i < cfg.num_dynamic_lights; ++i)
	{
		const uint32_t c = m.empty_cells[uint32_t(g() % m.empty_cells.size())];
		const uint32_t x = c % m.gx, y = (c / m.gx) % m.gy, z = c / (m.gx * m.gy);
		const float phase = uni(0.0f, 6.2831853f);
		const V3 pos = v3(CELL * (float(x) + 0.5f) + 40.0f * std::cos(cfg.time * 1.3f + phase),
			CELL * (float(y) + 0.5f) + 40.0f * std::sin(cfg.time * 1.1f + phase), CELL * (float(z) + 0.5f) + uni(-20.0f, 20.0f));
		const float radius = uni(200.0f, 450.0f);
		sw_light l;
		l.position[0] = pos.x;
		l.position[1] = pos.y;
		l.position[2] = pos.z;
		l.radius = radius;
		l.inv_square_radius = 1.0f / (radius * radius);
		l.color[0] = uni(0.3f, 1.0f) * 40000.0f;
		l.color[1] = uni(0.3f, 1.0f) * 40000.0f;
		l.color[2] = uni(0.3f, 1.0f) * 40000.0f;
		const uint32_t index = uint32_t(f.lights.size());
		f.lights.push_back(l);
		// leaves touched by the light's sphere (the reference's dynamic_objects_index)
		const int r = int(std::ceil(radius / CELL));
		for(int dz = -r; dz <= r; ++dz)
			for(int dy = -r; dy <= r; ++dy)
				for(int dx = -r; dx <= r; ++dx)
				{
					const int cx = int(x) + dx, cy = int(y) + dy, cz = int(z) + dz;
					if(cx < 0 || cy < 0 || cz < 0 || cx >= int(m.gx) || cy >= int(m.gy) || cz >= int(m.gz))
						continue;
					const uint32_t cell = cell_index(m, uint32_t(cx), uint32_t(cy), uint32_t(cz));
					if(m.cell[cell] != 0)
						continue;
					const BBox& b = m.leaf_bbox[m.cell_leaf[cell]];
					const float qx = std::max(b.min.x, std::min(pos.x, b.max.x)) - pos.x;
					const float qy = std::max(b.min.y, std::min(pos.y, b.max.y)) - pos.y;
					const float qz = std::max(b.min.z, std::min(pos.z, b.max.z)) - pos.z;
					if(qx * qx + qy * qy + qz * qz <= radius * radius)
						f.leaf_lights[m.cell_leaf[cell]].push_back(index);
				}
	}
}

void prepare_models(swh_frame& f, const swh_map& m, const swh_frame_config& cfg, V3 cam_pos, const M4& view, const M4& planes)
{
	(void)f;
	(void)m;
	(void)cfg;
	(void)cam_pos;
	(void)view;
	(void)planes;
}

void draw_leaf_objects(swh_frame& f, const swh_map& m, const swh_frame_config& cfg, uint32_t leaf, const V3 clip_planes[8],
	const int32_t clip_rect[4])
{
	(void)f;
	(void)m;
	(void)cfg;
	(void)leaf;
	(void)clip_planes;
	(void)clip_rect;
}

} // namespace swh
