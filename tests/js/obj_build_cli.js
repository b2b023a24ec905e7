'use strict'
// The demo's real flow on the GPU (main.js:72-88): OBJ text -> parseObj ->
// initBVHBuild(device).build(TRI_ARR, NUM_TRIS, BOUNDS).
//   node obj_build_cli.js <model.obj> <out_prefix>
// Writes <out_prefix>.bvh and .otri (the build's outputs) and, for the
// oracle, <out_prefix>.tri (TRI_ARR rounded to f32 as Float32Array.set
// rounds it) and <out_prefix>.bounds (BOUNDS rounded as DataView.setFloat32
// rounds it, min.xyz then max.xyz).
const fs = require('fs')
const path = require('path')
const bvh = require(path.join(__dirname, '..', '..', 'webgpu-lvbh-demo_amd', 'bvh'))
const { parseObj } = require(path.join(__dirname, '..', '..', 'webgpu-lvbh-demo_amd', 'bvh', 'obj.js'))

async function main () {
  const [objPath, prefix] = process.argv.slice(2)
  const { NUM_TRIS, TRI_ARR, BOUNDS } = parseObj(fs.readFileSync(objPath, 'utf8'))
  const device = bvh.createDevice({ ordinal: 0 })
  const BVH = bvh.initBVHBuild(device)
  const { BVH_BUFFER, O_TRIANGLE_BUFFER } = await BVH.build(TRI_ARR, NUM_TRIS, BOUNDS)
  fs.writeFileSync(prefix + '.bvh', Buffer.from(await BVH_BUFFER.read()))
  fs.writeFileSync(prefix + '.otri', Buffer.from(await O_TRIANGLE_BUFFER.read()))
  const t = new Float32Array(12 * NUM_TRIS)
  t.set(TRI_ARR)
  fs.writeFileSync(prefix + '.tri', Buffer.from(t.buffer))
  const b = new Float32Array([...BOUNDS.min, ...BOUNDS.max])
  fs.writeFileSync(prefix + '.bounds', Buffer.from(b.buffer))
  console.log(JSON.stringify({ ok: true, NUM_TRIS }))
}

main().catch((e) => { console.error(e); process.exit(1) })
